"""Build oracle/libpff_oracle.so from pff_oracle.c (gcc, OpenMP).

Test infrastructure: the oracle is the CPU restatement the HIP library is
checked against, and the CPU baseline timed by bench.py.  Built for
x86-64-v3 (AVX2: every host this build runs on has it; -march=native would
tie the library, which travels between hosts, to the build machine) and
without floating-point contraction, so its f32 arithmetic is the plain IEEE
one.
"""

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libpff_oracle.so")


def build(verbose: bool = False) -> str:
    src = os.path.join(HERE, "pff_oracle.c")
    tmp = LIB + ".tmp"
    cmd = ["gcc", "-O3", "-march=x86-64-v3", "-std=c11", "-fPIC", "-shared", "-fopenmp", "-ffp-contract=off", "-fno-fast-math",
           "-Wall", "-Wextra", "-Wno-unused-parameter", "-o", tmp, src, "-lm"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"oracle build failed:\n{r.stdout}\n{r.stderr}")
    if verbose and r.stderr.strip():
        print(r.stderr)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    print("built", LIB)
    sys.exit(0)
