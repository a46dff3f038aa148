"""Build recipe for the oracle (test infrastructure only).

Compiles oracle/phaser_oracle.c into oracle/libphaser_oracle.so with gcc and
-ffp-contract=off, so float expressions round as the reference's scalar C++
does.

build_ref() compiles the part of the reference's vendored SOFT library that
needs nothing beyond libc and libm -- makeWigner.c, makeweights.c,
s2_primitive.c and utils_so3.c from phaser_core/lib/SOFT, with the headers of
phaser_core/include -- where they lie under /root/reference, into
oracle/_ref/libsoft_ref.so, when that tree is present (the GPU box uses the
built file).  The tests compare the restatement's Wigner-d tables, quadrature
weights and coefficient index with it.  The rest of the reference's path
(SOFT's FFTW transforms, the Eigen/PCL/FLANN/glog/nlopt classes) needs headers
and libraries this image does not have and is not built; those functions are
pinned by the reference's own known-answer vectors, closed forms and brute
force instead (tests/).
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libphaser_oracle.so")


def build(verbose=False, src=None, out=None):
    """Compiles phaser_oracle.c (or the given source) into LIB (or out)."""
    src = src or os.path.join(HERE, "phaser_oracle.c")
    out = out or LIB
    tmp = out + f".tmp{os.getpid()}"  # concurrent builds (test workers) do not collide
    cmd = ["gcc", "-O2", "-std=gnu11", "-fPIC", "-shared", "-ffp-contract=off",
           "-fno-fast-math", "-Wall", "-Wno-unused-function", "-I", HERE, "-o", tmp, src, "-lm"]
    if verbose:
        print(" ".join(cmd))
    subprocess.check_call(cmd)
    os.replace(tmp, out)
    return out


REF_ROOT = "/root/reference"
REF_DIR = os.path.join(HERE, "_ref")
REF_LIB = os.path.join(REF_DIR, "libsoft_ref.so")
REF_SOURCES = ["makeWigner.c", "makeweights.c", "s2_primitive.c", "utils_so3.c"]


def build_ref(verbose=False):
    """oracle/_ref/libsoft_ref.so from the reference's own SOFT sources, or
    None when the reference tree is absent."""
    soft = os.path.join(REF_ROOT, "phaser_core", "lib", "SOFT")
    inc = os.path.join(REF_ROOT, "phaser_core", "include")
    if not os.path.isdir(soft):
        return None
    os.makedirs(REF_DIR, exist_ok=True)
    tmp = REF_LIB + ".tmp"
    cmd = ["gcc", "-O2", "-std=gnu11", "-fPIC", "-shared", "-I", inc, "-o", tmp,
           *[os.path.join(soft, f) for f in REF_SOURCES], "-lm"]
    if verbose:
        print(" ".join(cmd))
    subprocess.check_call(cmd)
    os.replace(tmp, REF_LIB)
    return REF_LIB


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    build_ref(verbose="-v" in sys.argv)
