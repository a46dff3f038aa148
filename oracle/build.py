"""Build recipe for the oracle (test infrastructure only).

Compiles oracle/phaser_oracle.c into oracle/libphaser_oracle.so with gcc and
-ffp-contract=off, so float expressions round as the reference's scalar C++
does.  The reference's own SOFT/FFTW path is not built here: it needs FFTW3,
Eigen, PCL, glog and nlopt headers, none of which exist in this image, so
oracle/_ref stays empty; the restatement is pinned by the reference's own
known-answer vectors, closed forms and brute force instead (tests/).
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libphaser_oracle.so")


def build(verbose=False):
    src = os.path.join(HERE, "phaser_oracle.c")
    tmp = LIB + ".tmp"
    cmd = ["gcc", "-O2", "-std=gnu11", "-fPIC", "-shared", "-ffp-contract=off",
           "-fno-fast-math", "-Wall", "-Wno-unused-function", "-o", tmp, src, "-lm"]
    if verbose:
        print(" ".join(cmd))
    subprocess.check_call(cmd)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
