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


# ---------------------------------------------------------------- spatial padding
def test_padded_size_truncates_like_the_reference(po):
    """spatial-correlation.cc:17-24 multiplies N^3 by ((N + 2 pad) / N)^3 in
    double and truncates to uint32: one less than (N + 2 pad)^3 for some
    sizes (values from the same double arithmetic, by hand)."""
    assert po.padded_size(3, 2) == 343
    assert po.padded_size(3, 4) == 1330   # 11^3 = 1331
    assert po.padded_size(5, 1) == 342    # 7^3 = 343
    assert po.padded_size(128, 16) == 160 ** 3
    assert po.padded_size(201, 10) == int(201 ** 3 * ((221 / 201.0) ** 3))


def _pad_index(i, N, pad):
    """computeZeroPaddedIndex (spatial-correlation.cc:223-230): Ind2Sub(i, N,
    N) = (x, y, z) with y = i mod N, x = (i / N) mod N, z = i / N^2
    (signal-utils.cc:11-19), then Sub2Ind(x + pad, y + pad, z + pad, Np, Np)
    = (x + pad) Np + (y + pad) + Np^2 (z + pad) (signal-utils.cc:5-9)."""
    Np = N + 2 * pad
    x, y, z = (i // N) % N, i % N, i // (N * N)
    return (x + pad) * Np + (y + pad) + Np * Np * (z + pad)


@pytest.mark.parametrize("N,pad,shifted,want_inside", [(12, 1, 1, True), (6, 2, 1, False),
                                                        (12, 1, 0, True), (9, 2, 1, True)])
def test_padded_inverse_of_one_product(po, N, pad, shifted, want_inside):
    """One nonzero product v at index i0: the padded buffer holds v at
    p0 = computeZeroPaddedIndex(i0); the Laplace class's IFFTShift of the
    buffer's first N^3 entries (rotate left by N^3 / 2, signal-utils.h:39-47)
    moves it to q0 = (p0 - N^3 / 2) mod N^3 when p0 < N^3; the c2r of size Np
    reads q0 as (k0, k1, k2) of an Np x Np x (Np/2 + 1) half spectrum and,
    for 0 < k2 < Np/2, returns the closed form 2 Re(v e^{2 pi i k.n / Np})."""
    T, Np = N ** 3, N + 2 * pad
    hp = Np // 2 + 1
    v = 0.7 - 0.3j
    found = None
    for i0 in range(T):
        p0 = _pad_index(i0, N, pad)
        inside = p0 < T
        if inside != want_inside:
            continue
        q0 = (p0 - T // 2) % T if (shifted and inside) else p0
        k0, k1, k2 = q0 // (Np * hp), (q0 // hp) % Np, q0 % hp
        if q0 < Np * Np * hp and 0 < k2 and 2 * k2 < Np:
            found = (i0, k0, k1, k2)
            break
    assert found is not None
    i0, k0, k1, k2 = found
    prod = np.zeros(T, np.complex128)
    prod[i0] = v
    got = po.spatial_pad_inverse(prod, N, pad, shifted)
    n0, n1, n2 = np.meshgrid(np.arange(Np), np.arange(Np), np.arange(Np), indexing="ij")
    expect = 2.0 * np.real(v * np.exp(2j * np.pi * (k0 * n0 + k1 * n1 + k2 * n2) / Np))
    np.testing.assert_allclose(got, expect.reshape(-1), atol=1e-12)


@pytest.mark.parametrize("N,pad", [(8, 1), (7, 2)])
def test_padded_spatial_fusion_composition(po, N, pad):
    """SpatialCorrelationLaplace with zero padding (spatial-correlation-
    laplace.cc:17-51, spatial-correlation.cc:142-161,221-230): the fused
    product of the shifted spectra, placed by computeZeroPaddedIndex,
    IFFTShift of the first N^3 entries, c2r of size N + 2 pad -- composed
    from numpy's rfftn / fftshift / ifftshift / irfftn and the pinned fusion."""
    rng = np.random.default_rng(70 + N)
    C = 4
    f = rng.normal(size=(C, N, N, N))
    g = rng.normal(size=(C, N, N, N))
    f[1] *= 3.0
    g[2] *= 3.0
    T, h, Np = N ** 3, N // 2 + 1, N + 2 * pad
    hp = Np // 2 + 1

    def spectra(x):
        out = np.zeros((C, T), np.complex128)
        for c in range(C):
            out[c, : N * N * h] = np.fft.rfftn(x[c]).reshape(-1)
        return out

    F = po.fuse_channels(np.stack([np.fft.fftshift(s) for s in spectra(f)]), levels=5, div=4.0)
    G = po.fuse_channels(np.stack([np.fft.fftshift(s) for s in spectra(g)]), levels=5, div=4.0)
    Cpad = np.zeros(Np ** 3, np.complex128)
    Cpad[_pad_index(np.arange(T), N, pad)] = F * np.conj(G)
    Cpad[:T] = np.fft.ifftshift(Cpad[:T])
    expect = np.fft.irfftn(Cpad[: Np * Np * hp].reshape(Np, Np, hp), s=(Np, Np, Np),
                           axes=(0, 1, 2)) * Np ** 3
    got = po.spatial_correlate_laplace(f.reshape(C, -1), g.reshape(C, -1), N, pad=pad)
    np.testing.assert_allclose(got, expect.reshape(-1), atol=1e-10 * np.max(np.abs(expect)))


# ---------------------------------------------------------------- low-pass classes
def test_spherical_low_pass_full_band_is_the_plain_correlation(po):
    """Bounds covering all B^2 coefficients (the flags' defaults 0 / 10000)
    leave the correlation of the first grid of each signal unchanged."""
    rng = np.random.default_rng(5)
    bw = 6
    n = 2 * bw
    f, h = rng.normal(size=(n, n)), rng.normal(size=(n, n))
    got = po.spherical_correlate_low_pass(bw, f, h, 0, 10000)
    np.testing.assert_allclose(got, po.so3_correlate(bw, po.fst(bw, f), po.fst(bw, h)),
                               rtol=1e-12, atol=1e-14)


def test_spherical_low_pass_keeps_the_zonal_coefficients(po):
    """Seanindex j sits at FFTShift position (j - B^2 / 2) mod B^2 (B^2
    even): the band [B^2 / 2, B^2 / 2 + B) keeps j < B, the order m = 0 of
    every degree.  The combine then pairs f_{l,0} with h_{l,0} only
    (so3_correlate_fftw.c:120-157), so |C|^2 is constant along alpha and
    gamma and, along beta, proportional to (sum_l 2 pi sqrt(2/(2l+1))
    f_{l,0} h_{l,0} d~^l_{0,0}(beta))^2 with the pinned Wigner-d; an empty
    band gives zero."""
    rng = np.random.default_rng(6)
    bw = 6
    n, full = 2 * bw, bw * bw
    f, h = rng.normal(size=(n, n)), rng.normal(size=(n, n))
    got = po.spherical_correlate_low_pass(bw, f, h, full // 2, full // 2 + bw)
    got = got.reshape(n, n, n)  # [beta][alpha][gamma]
    np.testing.assert_allclose(got, got[:, :1, :1] * np.ones((1, n, n)), rtol=1e-9, atol=1e-12)
    fc, hc = po.fst(bw, f), po.fst(bw, h)
    d = po.wigner_d(bw, 0, 0)  # [l][beta_k]
    w = np.array([2 * np.pi * np.sqrt(2.0 / (2 * l + 1)) for l in range(bw)])
    amp = (w * fc[:bw].real * hc[:bw].real) @ d
    np.testing.assert_allclose(got[:, 0, 0] / got[:, 0, 0].max(), amp ** 2 / np.max(amp ** 2),
                               atol=1e-9)
    assert np.max(po.spherical_correlate_low_pass(bw, f, h, 5, 5)) == 0.0


def test_spatial_low_pass_full_cube_is_the_plain_correlation(po):
    """The whole cube [0, N-1]^3 multiplies every shifted index: shift,
    product, unshift is the plain product, so the result is
    SpatialCorrelation's (pinned by the reference's known answers)."""
    rng = np.random.default_rng(8)
    N = 6
    f, g = rng.normal(size=N ** 3), rng.normal(size=N ** 3)
    np.testing.assert_allclose(po.spatial_correlate_low_pass(f, g, N, 0, N - 1),
                               po.spatial_correlate(f, g, N), atol=1e-10)
    with pytest.raises(po.PeakWindowError):
        po.spatial_correlate_low_pass(f, g, N, 0, N)  # the capped bound N reaches N^3


def test_spatial_low_pass_one_index(po):
    """A cube of one index keeps the product at shifted index i0 = Sub2Ind(
    b, b, b, N, N) = b N + b + N^2 b, i.e. the spectra's entry u0 =
    (i0 + ceil(N^3 / 2)) mod N^3 of the r2c buffer: the padded inverse of
    that one product (the pinned closed form)."""
    rng = np.random.default_rng(9)
    N, b, pad = 8, 5, 1
    f, g = rng.normal(size=N ** 3), rng.normal(size=N ** 3)
    i0 = b * N + b + N * N * b
    u0 = (i0 + N ** 3 // 2) % N ** 3
    F, G = po.r2c3(f), po.r2c3(g)
    P = np.zeros(N ** 3, np.complex128)
    P[i0] = F[u0] * np.conj(G[u0])
    np.testing.assert_allclose(po.spatial_correlate_low_pass(f, g, N, b, b, pad=pad),
                               po.spatial_pad_inverse(P, N, pad, True), atol=1e-10)
