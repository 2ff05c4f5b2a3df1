"""Learned PPE preconditioner: binding of the vox_lpc_* entry points.

The network, its training and its use inside flexible CG run in
libvoxarmps.so (csrc/vox_lpc.hip, csrc/vox_cg.hip); this class only holds the
handle and moves arrays.
"""
import ctypes as C

import numpy as np

from . import _abi
from .sim import hip_backend

NPARAM = _abi.LPC_NPARAM


def _fptr(a):
    return None if a is None else a.ctypes.data_as(C.POINTER(C.c_float))


class LearnedPreconditioner:
    """Network weights, optimiser state and a training set of PPE systems,
    resident on the GPU."""

    def __init__(self, seed=0, backend=None):
        self.be = backend or hip_backend()
        self.h = C.c_void_p()
        self.be.check(self.be.lib.vox_lpc_create(C.byref(self.h), C.c_uint64(seed)), "lpc_create")

    def close(self):
        if self.h:
            self.be.lib.vox_lpc_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def weights(self):
        w = np.empty(NPARAM, np.float32)
        self.be.check(self.be.lib.vox_lpc_get_weights(self.h, _fptr(w)), "lpc_get_weights")
        return w

    def set_weights(self, w):
        w = np.ascontiguousarray(w, np.float32)
        if w.shape != (NPARAM,):
            raise ValueError(f"expected {NPARAM} weights")
        self.be.check(self.be.lib.vox_lpc_set_weights(self.h, _fptr(w)), "lpc_set_weights")

    def dump(self, sim):
        """Store the simulation's assembled PPE as a training system; returns its index."""
        k = C.c_int32()
        self.be.check(self.be.lib.vox_lpc_dump(self.h, sim.ctx, C.byref(k)), "lpc_dump")
        return k.value

    def count(self):
        n, rows = C.c_int32(), C.c_int64()
        self.be.check(self.be.lib.vox_lpc_count(self.h, C.byref(n), C.byref(rows)), "lpc_count")
        return n.value, rows.value

    def eval(self, index, nu):
        """Features (nu x 8), g and c of stored system `index` (nu rows)."""
        f = np.empty((nu, 8), np.float32)
        g = np.empty(nu, np.float32)
        c = np.empty(nu, np.float32)
        self.be.check(self.be.lib.vox_lpc_eval(self.h, index, _fptr(f), _fptr(g), _fptr(c)),
                      "lpc_eval")
        return f, g, c

    def loss_grad(self, index, seed):
        loss = C.c_double()
        g = np.empty(NPARAM, np.float32)
        self.be.check(self.be.lib.vox_lpc_loss_grad(self.h, index, C.c_uint64(seed), C.byref(loss),
                                                    _fptr(g)), "lpc_loss_grad")
        return loss.value, g

    def train(self, epochs, lr=1e-3, seed=0):
        """Adam over every stored system; returns the mean loss of the first
        and the last pass."""
        a, b = C.c_double(), C.c_double()
        self.be.check(self.be.lib.vox_lpc_train(self.h, int(epochs), float(lr), C.c_uint64(seed),
                                                C.byref(a), C.byref(b)), "lpc_train")
        return a.value, b.value
