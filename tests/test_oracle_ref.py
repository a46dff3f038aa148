"""The oracle against the reference's own code where it builds here:
oracle/_ref/libsoft_ref.so holds the reference's vendored SOFT sources that
need only libc/libm (makeWigner.c, makeweights.c, s2_primitive.c,
utils_so3.c; compiled by oracle/build.py's build_ref from the reference
tree).  Compared: the L2-normalised Wigner-d tables of every order pair on
the beta grid (genWig_L2 with SOFT's own evaluation points) against
po_wigner_d, the quadrature weights (makeweights) against po_makeweights, and
the coefficient index (seanindex) against po_seanindex."""
import ctypes as C
import os

import numpy as np
import pytest

from oracle import build as ob
from oracle import oracle

D = C.POINTER(C.c_double)


@pytest.fixture(scope="module")
def soft():
    if not os.path.exists(ob.REF_LIB):
        pytest.skip("oracle/_ref is not built (the reference tree is absent)")
    L = C.CDLL(ob.REF_LIB)
    for name in ("SinEvalPts", "CosEvalPts", "SinEvalPts2", "CosEvalPts2"):
        getattr(L, name).argtypes = [C.c_int, D]
    L.genWig_L2.argtypes = [C.c_int, C.c_int, C.c_int, D, D, D, D, D, D]
    L.makeweights.argtypes = [C.c_int, D]
    L.seanindex.argtypes = [C.c_int, C.c_int, C.c_int]
    L.seanindex.restype = C.c_int
    return L


def _p(a):
    return a.ctypes.data_as(D)


@pytest.mark.parametrize("bw", [4, 8, 16])
def test_wigner_d_matches_soft(soft, bw):
    n = 2 * bw
    se, ce, se2, ce2 = (np.zeros(n) for _ in range(4))
    soft.SinEvalPts(n, _p(se))
    soft.CosEvalPts(n, _p(ce))
    soft.SinEvalPts2(n, _p(se2))
    soft.CosEvalPts2(n, _p(ce2))
    work = np.zeros(6 * n)
    worst = 0.0
    for m1 in range(-(bw - 1), bw):
        for m2 in range(-(bw - 1), bw):
            m = max(abs(m1), abs(m2))
            ref = np.zeros((bw - m) * n)
            soft.genWig_L2(m1, m2, bw, _p(se), _p(ce), _p(se2), _p(ce2), _p(ref), _p(work))
            got = oracle.wigner_d(bw, m1, m2).ravel()
            worst = max(worst, float(np.max(np.abs(got - ref))))
    assert worst < 1e-12, worst


@pytest.mark.parametrize("bw", [2, 8, 32, 64])
def test_weights_and_index_match_soft(soft, bw):
    ref = np.zeros(4 * bw)
    soft.makeweights(bw, _p(ref))
    np.testing.assert_allclose(oracle.makeweights(bw), ref, rtol=1e-13, atol=1e-16)
    for l in range(bw):
        for m in range(-l, l + 1):
            assert oracle.seanindex(m, l, bw) == soft.seanindex(m, l, bw)
