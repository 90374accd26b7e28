"""Compile the C oracle (oracle/fdtd_oracle.c) into oracle/liboracle_fdtd.so.

Test infrastructure: gcc -O3 -fopenmp with contraction off, so its float32
operation order equals the HIP kernels' and the reference's expressions.
"""

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "liboracle_fdtd.so")


def build(verbose: bool = False) -> list:
    cmd = ["gcc", "-O3", "-march=x86-64-v3", "-fopenmp", "-ffp-contract=off", "-fPIC", "-shared",
           "-std=c11", os.path.join(HERE, "fdtd_oracle.c"), "-o", OUT + ".tmp", "-lm"]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("gcc failed for the C oracle")
    os.replace(OUT + ".tmp", OUT)
    return [OUT]


if __name__ == "__main__":
    print(build(verbose=True))
