"""Pins the oracle's compositions that the per-function tests leave open:
the SO(3) correlation's per-degree combine weight and global scale, and the
coefficient shifts around both Laplace fusions.

* SO(3): the mathematics fixes the whole complex grid.  For band-limited
  real f and h the reference's correlation at bin (k, j, i) is the integral
  over the sphere of f(w) h(R^-1 w) with R = Rz(alpha_j) Ry(beta_k)
  Rz(gamma_i), alpha_j = 2 pi j / 2B, beta_k = pi (2k+1) / 4B,
  gamma_i = 2 pi i / 2B (the ZYZ decode of rotation-utils.cc:197-207): the
  combine weight 2 pi sqrt(2/(2l+1)) (so3_correlate_fftw.c:120) cancels the
  L2 normalisation of the Wigner-d tables up to 2 pi, and the scale
  (1/2B)(B/pi) of soft_fftw.c:1905-1911 removes that 2 pi, so the constant
  is exactly 1.  The tests evaluate the integral by Gauss-Legendre x uniform
  quadrature (exact for the product's degree) of scipy's spherical
  harmonics, independently of the oracle's Wigner recurrences, combine and
  FFTs; a wrong per-degree weight, sign, m-dependence or layout fails them.
* Spherical fusion (spherical-correlation-laplace.cc:23-31, 101-113): the
  fused coefficients are ifftshift(fuse(fftshift(FST(channel)))) over the
  flat seanindex array, composed here from numpy's fftshift/ifftshift and
  the fusion already pinned by the reference's pyramid known answers.
* Spatial fusion (spatial-correlation-laplace.cc:17-89): r2c into an N^3
  buffer with a zero tail, 1-D fftshift of the N^3 array, fusion, the
  unnormalised cross spectrum, ifftshift and c2r of the first N^2(N/2+1)
  entries, composed from numpy's rfftn / fftshift / irfftn and the pinned
  fusion.
Random inputs are chosen so that the shifted and unshifted band
assignments differ (a dropped shift changes the fused result)."""
import numpy as np
import pytest
import scipy.special as sp


# ---------------------------------------------------------------- SO(3)
def _real_coefficients(po, bw, rng):
    """Coefficients of a real band-limited function: f_{l,-m} =
    (-1)^m conj(f_{l,m}) (orthonormal Y with the Condon-Shortley phase)."""
    c = np.zeros(bw * bw, np.complex128)
    for l in range(bw):
        c[po.seanindex(0, l, bw)] = rng.normal()
        for m in range(1, l + 1):
            v = rng.normal() + 1j * rng.normal()
            c[po.seanindex(m, l, bw)] = v
            c[po.seanindex(-m, l, bw)] = (-1) ** m * np.conj(v)
    return c


def _evaluate(po, bw, coef, theta, phi):
    out = np.zeros(np.shape(theta), np.complex128)
    for l in range(bw):
        for m in range(-l, l + 1):
            out += coef[po.seanindex(m, l, bw)] * sp.sph_harm_y(l, m, theta, phi)
    return out.real


def _rz(a):
    c, s = np.cos(a), np.sin(a)
    return np.array([[c, -s, 0.0], [s, c, 0.0], [0.0, 0.0, 1.0]])


def _ry(a):
    c, s = np.cos(a), np.sin(a)
    return np.array([[c, 0.0, s], [0.0, 1.0, 0.0], [-s, 0.0, c]])


def _brute_correlation(po, bw, f, h, bins):
    """integral of f(w) h(R^-1 w) over the sphere at each (k, j, i) bin."""
    n = 2 * bw
    L, M = 2 * bw + 2, 4 * bw + 4  # exact for the product (degree <= 2bw - 2)
    x, wx = np.polynomial.legendre.leggauss(L)
    th = np.arccos(x)
    ph = 2 * np.pi * np.arange(M) / M
    TH, PH = np.meshgrid(th, ph, indexing="ij")
    W = np.repeat(wx[:, None], M, 1) * (2 * np.pi / M)
    V = np.stack([np.sin(TH) * np.cos(PH), np.sin(TH) * np.sin(PH), np.cos(TH)], -1)
    fv = _evaluate(po, bw, f, TH, PH)
    out = []
    for k, j, i in bins:
        R = _rz(2 * np.pi * j / n) @ _ry(np.pi * (2 * k + 1) / (4 * bw)) @ _rz(2 * np.pi * i / n)
        U = V @ R  # rows R^T w = R^-1 w
        t = np.arccos(np.clip(U[..., 2], -1.0, 1.0))
        p = np.arctan2(U[..., 1], U[..., 0])
        out.append(np.sum(W * fv * _evaluate(po, bw, h, t, p)))
    return np.array(out)


@pytest.mark.parametrize("bw", [3, 4, 6])
def test_so3_correlation_is_the_rotation_integral(po, bw):
    rng = np.random.default_rng(100 + bw)
    f = _real_coefficients(po, bw, rng)
    h = _real_coefficients(po, bw, rng)
    n = 2 * bw
    mag, cp = po.so3_correlate(bw, f, h, want_complex=True)
    if bw <= 4:
        bins = list(np.ndindex(n, n, n))
    else:
        bins = [tuple(b) for b in rng.integers(0, n, size=(96, 3))]
    brute = _brute_correlation(po, bw, f, h, bins)
    flat = [(k * n + j) * n + i for k, j, i in bins]
    scale = np.max(np.abs(brute))
    np.testing.assert_allclose(cp[flat].real, brute, atol=1e-10 * scale)
    np.testing.assert_allclose(cp[flat].imag, 0.0, atol=1e-10 * scale)
    np.testing.assert_allclose(mag[flat], brute ** 2, atol=1e-10 * scale ** 2)


@pytest.mark.parametrize("l1,l2", [(0, 3), (1, 5), (2, 7)])
def test_so3_correlation_of_zonal_functions(po, l1, l2):
    """f = h = Y_l1^0 + Y_l2^0: the correlation is P_l1(cos beta) +
    P_l2(cos beta) at every alpha and gamma (Legendre addition theorem),
    so the two degrees' terms fix the weight's l-dependence and the scale."""
    bw = 8
    n = 2 * bw
    f = np.zeros(bw * bw, np.complex128)
    f[po.seanindex(0, l1, bw)] = 1.0
    f[po.seanindex(0, l2, bw)] = 1.0
    mag, cp = po.so3_correlate(bw, f, f, want_complex=True)
    beta = np.pi * (2 * np.arange(n) + 1) / (4 * bw)
    expect = sp.eval_legendre(l1, np.cos(beta)) + sp.eval_legendre(l2, np.cos(beta))
    grid = cp.reshape(n, n, n)
    np.testing.assert_allclose(grid.real, np.broadcast_to(expect[:, None, None], grid.shape),
                               atol=1e-12)
    np.testing.assert_allclose(grid.imag, 0.0, atol=1e-12)


# ---------------------------------------------------------------- fusion shifts
def _random_grids(rng, C, n):
    g = rng.normal(size=(C, n * n))
    g[0] *= 3.0  # channels of different energy, so the max picks vary by band
    return g


@pytest.mark.parametrize("bw", [7, 8])
def test_spherical_fusion_is_shift_fuse_unshift(po, bw):
    rng = np.random.default_rng(bw)
    n = 2 * bw
    f = _random_grids(rng, 4, n)
    h = _random_grids(rng, 4, n)
    mag, ff, hf = po.spherical_correlate_laplace(bw, f, h, fused=True)
    for grids, got in ((f, ff), (h, hf)):
        coefs = np.stack([po.fst(bw, g.reshape(n, n)) for g in grids])
        shifted = np.stack([np.fft.fftshift(c) for c in coefs])
        expect = np.fft.ifftshift(po.fuse_channels(shifted, levels=5, div=4.0))
        np.testing.assert_allclose(got, expect, atol=1e-12)
        # the shift matters for this input: fusing unshifted differs
        assert np.max(np.abs(po.fuse_channels(coefs, levels=5, div=4.0) - expect)) > 1e-3
    np.testing.assert_allclose(mag, po.so3_correlate(bw, ff, hf), rtol=1e-12, atol=1e-14)


@pytest.mark.parametrize("N", [5, 8])
def test_spatial_fusion_is_shift_fuse_unshift(po, N):
    rng = np.random.default_rng(40 + N)
    C = 4
    f = rng.normal(size=(C, N, N, N))
    g = rng.normal(size=(C, N, N, N))
    f[1] *= 3.0
    g[2] *= 3.0
    T, h = N ** 3, N // 2 + 1

    def spectra(x):
        out = np.zeros((C, T), np.complex128)
        for c in range(C):
            out[c, : N * N * h] = np.fft.rfftn(x[c]).reshape(-1)  # zero tail
        return out

    Fs, Gs = spectra(f), spectra(g)
    F = po.fuse_channels(np.stack([np.fft.fftshift(s) for s in Fs]), levels=5, div=4.0)
    G = po.fuse_channels(np.stack([np.fft.fftshift(s) for s in Gs]), levels=5, div=4.0)
    Cs = np.fft.ifftshift(F * np.conj(G))
    expect = np.fft.irfftn(Cs[: N * N * h].reshape(N, N, h), s=(N, N, N), axes=(0, 1, 2)) * T
    got = po.spatial_correlate_laplace(f.reshape(C, -1), g.reshape(C, -1), N)
    scale = np.max(np.abs(expect))
    np.testing.assert_allclose(got, expect.reshape(-1), atol=1e-10 * scale)
    # the shift matters for this input: fusing the unshifted spectra differs
    F0 = po.fuse_channels(Fs, levels=5, div=4.0)
    G0 = po.fuse_channels(Gs, levels=5, div=4.0)
    C0 = F0 * np.conj(G0)
    other = np.fft.irfftn(C0[: N * N * h].reshape(N, N, h), s=(N, N, N), axes=(0, 1, 2)) * T
    assert np.max(np.abs(other - expect)) > 1e-3 * scale
