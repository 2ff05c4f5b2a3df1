"""Build the CPU oracle (test infrastructure) as oracle/liboracle.so.

gcc without -mfma and with -ffp-contract=off keeps every double operation a
separate IEEE operation in source order, like the HIP library. -fopenmp runs
the per-particle loops on the host's cores (OMP_NUM_THREADS); each particle's
sums keep their order and the reductions left in parallel loops are max and
integer counts, so results do not depend on the thread count.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")


def build_oracle(verbose=False):
    tmp = LIB + ".tmp"
    cmd = ["gcc", "-O2", "-std=c99", "-fPIC", "-shared", "-ffp-contract=off", "-fno-fast-math",
           "-fopenmp", "-Wall", "-o", tmp, os.path.join(HERE, "oracle.c"), "-lm"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"oracle build failed:\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}", file=sys.stderr)
    return LIB


if __name__ == "__main__":
    build_oracle(verbose=True)
