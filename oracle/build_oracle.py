"""Build recipe for the oracle library (test infrastructure only).

    python oracle/build_oracle.py

compiles oracle/edm_oracle.c with gcc into oracle/_build/libedm_oracle.so.
-ffp-contract=off keeps every multiply and add separately rounded, as the
reference's numba kernels are (numba does not enable fast-math contraction).
There is no oracle/_ref build: the reference is Python + numba and ships no
C/C++ sources to compile.
"""

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "edm_oracle.c")
OUT_DIR = os.path.join(HERE, "_build")
OUT = os.path.join(OUT_DIR, "libedm_oracle.so")


def build(verbose: bool = False) -> str:
    os.makedirs(OUT_DIR, exist_ok=True)
    cmd = ["gcc", "-O2", "-fPIC", "-shared", "-std=c11", "-ffp-contract=off",
           "-fno-fast-math", "-fopenmp", "-o", OUT + ".tmp", SRC, "-lm"]
    if verbose:
        print(" ".join(cmd))
    subprocess.run(cmd, check=True)
    os.replace(OUT + ".tmp", OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
