"""Build libvoxarmps.so in-tree with hipcc for gfx950.

Every source is compiled with -ffp-contract=off (no FMA contraction, so FP64
per-particle sums follow the reference's operation order exactly), except the
FP32 stages of the mixed-precision mode, and with packed-FP32 instructions
disabled.
"""
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libvoxarmps.so")
CLI = os.path.join(HERE, "voxmps")
SOURCES = ["vox_api.hip", "vox_neighbor.hip", "vox_stages.hip", "vox_stages32.hip", "vox_ppe.hip",
           "vox_cgm.hip",
           "vox_cg.hip", "vox_lpc.hip", "vox_io.cpp"]
# The FP32 stages of the mixed-precision mode have no bit-exact counterpart, so
# they let the compiler fuse multiply-adds (the last -ffp-contract wins).
EXTRA = {"vox_stages32.hip": ["-ffp-contract=fast"]}
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def _compile(src, build_dir=BUILD, defines=()):
    obj = os.path.join(build_dir, os.path.basename(src) + ".o")
    cmd = [HIPCC, *FLAGS, *EXTRA.get(src, []), *[f"-D{d}" for d in defines], "-x", "hip", "-c",
           os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr}")
    return obj


def build_library(verbose=False, defines=(), out=LIB):
    """Compile every HIP/C++ source and link libvoxarmps.so (always rebuilds).

    defines: extra NAME=VALUE macros, e.g. the gather chunk sizes VOX_ASM_CHUNK,
    VOX_PRED_CHUNK, VOX_SPMV_CHUNK, VOX_LIST_CHUNK; out: library path."""
    build_dir = BUILD if out == LIB else out + ".build"
    os.makedirs(build_dir, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=min(len(SOURCES), os.cpu_count() or 4)) as ex:
        objs = list(ex.map(lambda s: _compile(s, build_dir, defines), SOURCES))
    LIBOUT = out
    tmp = LIBOUT + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, LIBOUT)
    if verbose:
        print(f"built {LIBOUT}", file=sys.stderr)
    if out == LIB:
        build_cli(verbose)
    return LIBOUT


def build_cli(verbose=False):
    """voxmps, the reference's 16-argument command line over the library
    (host C++ only, linked against libvoxarmps.so next to it)."""
    tmp = CLI + ".tmp"
    cmd = ["g++", "-std=c++17", "-O2", "-Wall", os.path.join(CSRC, "vox_cli.cpp"), "-o", tmp,
           "-L" + HERE, "-lvoxarmps", "-Wl,-rpath,$ORIGIN"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"voxmps build failed:\n{r.stderr}")
    os.replace(tmp, CLI)
    if verbose:
        print(f"built {CLI}", file=sys.stderr)
    return CLI


if __name__ == "__main__":
    build_library(verbose=True)
