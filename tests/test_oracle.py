"""Pins the FP64 oracle: reference known-answer tests, and mathematics that
fixes each function independently of the oracle's own formulas."""
import math

import numpy as np
import pytest
import scipy.special as sp


# ---------------------------------------------------------------- FFT
@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 6, 7, 8, 12, 48, 64, 96, 128, 192, 256])
def test_fft_matches_numpy(po, n):
    rng = np.random.default_rng(n)
    x = rng.normal(size=n) + 1j * rng.normal(size=n)
    np.testing.assert_allclose(po.fft(x, -1), np.fft.fft(x), rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(po.fft(x, +1), np.fft.ifft(x) * n, rtol=1e-12, atol=1e-12)


# ---------------------------------------------------------------- S^2
def test_seanindex_layout(po):
    bw = 9
    idx = {}
    for m in range(-(bw - 1), bw):
        for l in range(abs(m), bw):
            idx[(m, l)] = po.seanindex(m, l, bw)
    assert sorted(idx.values()) == list(range(bw * bw))
    # s2_semi_memo.c:100-118 ordering: m=0 row first, then m=1.., then m=-(bw-1)..-1
    assert idx[(0, 0)] == 0 and idx[(0, bw - 1)] == bw - 1 and idx[(1, 1)] == bw
    assert idx[(-(bw - 1), bw - 1)] == bw * (bw + 1) // 2
    assert idx[(-1, bw - 1)] == bw * bw - 1


@pytest.mark.parametrize("bw", [4, 16, 32])
def test_weights_integrate_polynomials_exactly(po, bw):
    w = po.makeweights(bw)[: 2 * bw]
    th = np.pi * (2 * np.arange(2 * bw) + 1) / (4 * bw)
    assert abs(w.sum() - 2.0) < 1e-13
    x = np.cos(th)
    for deg in range(0, 2 * bw):  # DH quadrature is exact below degree 2bw
        exact = 0.0 if deg % 2 else 2.0 / (deg + 1)
        assert abs(np.dot(w, x ** deg) - exact) < 1e-12


def _legendre_ref(l, m, x):
    """L2-normalised over [-1,1], Condon-Shortley phase (scipy lpmv has it)."""
    norm = math.sqrt((2 * l + 1) / 2 * math.factorial(l - m) / math.factorial(l + m))
    return norm * sp.lpmv(m, l, x)


@pytest.mark.parametrize("m", [0, 1, 2, 5, 11])
def test_legendre_against_scipy(po, m):
    bw = 16
    P = po.legendre(bw, m)
    x = np.cos(np.pi * (2 * np.arange(2 * bw) + 1) / (4 * bw))
    for l in range(m, bw):
        np.testing.assert_allclose(P[l - m], _legendre_ref(l, m, x), rtol=1e-10, atol=1e-12)


def _dh_grid(bw):
    n = 2 * bw
    th = np.pi * (2 * np.arange(n) + 1) / (4 * bw)
    ph = 2 * np.pi * np.arange(n) / n
    return np.meshgrid(th, ph, indexing="ij")


@pytest.mark.parametrize("bw", [8, 16])
def test_fst_recovers_spherical_harmonics(po, bw):
    """Samples of sum_k a_k Re(Y_lk^mk) must return coefficients a_k/2 at
    (l,m) and (-1)^m a_k/2 at (l,-m); scipy's sph_harm_y is orthonormal with
    the Condon-Shortley phase."""
    TH, PH = _dh_grid(bw)
    rng = np.random.default_rng(bw)
    terms = [(0, 0), (1, 1), (3, 2), (bw - 1, bw - 1), (bw - 1, 0), (bw // 2, 3)]
    amps = rng.normal(size=len(terms))
    f = np.zeros_like(TH)
    expect = np.zeros(bw * bw, np.complex128)
    for (l, m), a in zip(terms, amps):
        f += a * np.real(sp.sph_harm_y(l, m, TH, PH))
        if m == 0:
            expect[po.seanindex(0, l, bw)] += a
        else:
            expect[po.seanindex(m, l, bw)] += a / 2
            expect[po.seanindex(-m, l, bw)] += (-1) ** m * a / 2
    got = po.fst(bw, f)
    np.testing.assert_allclose(got, expect, atol=1e-11)


def test_fst_imaginary_part(po):
    bw = 8
    TH, PH = _dh_grid(bw)
    f = np.imag(sp.sph_harm_y(5, 3, TH, PH))  # = (Y53 - conj Y53)/(2i)
    got = po.fst(bw, f)
    expect = np.zeros(bw * bw, np.complex128)
    expect[po.seanindex(3, 5, bw)] = 1 / (2j)
    expect[po.seanindex(-3, 5, bw)] = -((-1) ** 3) / (2j)
    np.testing.assert_allclose(got, expect, atol=1e-12)


# ---------------------------------------------------------------- fusion
def test_pyramid_struct_kat(po):
    # pyramid-struct-test.cc:21-33
    coeffs, lower, upper = po.pyramid_struct(100, 2, 4.0)
    assert list(coeffs) == [50, 25]
    assert list(lower) == [25, 13]
    assert list(upper) == [75, 38]


def test_laplace_max_coeff_kat(po):
    # laplace-pyramid-test.cc:98-127: channels of 1 and 15, one level: the
    # fused laplacian outside [2,6) takes the larger channel.
    ch = np.stack([np.full(8, 1 + 1j), np.full(8, 15 + 15j)])
    fused = po.fuse_channels(ch, levels=1)
    for i in (0, 1, 6, 7):
        assert fused[i] == 15 + 15j


def test_laplace_low_pass_average_kat(po):
    # laplace-pyramid-test.cc:129-149: low pass of 5 and 15 averages to 10
    ch = np.stack([np.full(8, 5 + 5j), np.full(8, 15 + 15j)])
    fused = po.fuse_channels(ch, levels=1)
    for i in range(2, 6):
        assert fused[i] == 10 + 10j


@pytest.mark.parametrize("n,levels", [(8, 2), (80, 3)])
def test_laplace_fuse_positive_kat(po, n, levels):
    # laplace-pyramid-test.cc:151-187
    ch = np.stack([np.full(n, 5 + 5j), np.full(n, 15 + 15j)])
    fused = po.fuse_channels(ch, levels=levels)
    assert np.all(fused.real > 0) and np.all(fused.imag > 0)


def test_fusion_first_channel_wins_ties(po):
    ch = np.stack([np.full(64, 3 + 4j), np.full(64, -4 + 3j), np.full(64, 5 + 0j)])
    fused = po.fuse_channels(ch, levels=1)
    assert fused[0] == 3 + 4j  # equal energies: first channel


@pytest.mark.parametrize("n", [4096, 1024, 2304, 100, 50, 6, 8192 + 2])
def test_fusion_modes_reproduce_literal_fusion(po, n):
    """pick/mean/zero per index, derived once, must reproduce the literal
    pyramid on random data (mean taken over all channels)."""
    rng = np.random.default_rng(n)
    C = 4
    ch = rng.normal(size=(C, n)) + 1j * rng.normal(size=(C, n))
    ch[2:] = 0
    modes = po.fusion_modes(n)
    lit = po.fuse_channels(ch)
    e = np.abs(ch) ** 2
    pick = ch[np.argmax(e, axis=0), np.arange(n)]
    mean = ch.sum(axis=0) / C
    want = np.where(modes == 0, pick, np.where(modes == 1, mean, 0))
    np.testing.assert_allclose(lit, want, rtol=1e-15, atol=1e-15)


def test_fftshift_kats(po):
    # signal-utils-test.cc:32-110: even n swaps halves, odd n rotates by n//2+1
    x = np.arange(10) + 0j
    s = po.fftshift(x)
    np.testing.assert_array_equal(s[5:], x[:5])
    np.testing.assert_array_equal(s[:5], x[5:])
    x7 = np.arange(7) + 0j
    s7 = po.fftshift(x7)
    np.testing.assert_array_equal(s7[3:], x7[:4])
    np.testing.assert_array_equal(s7[:3], x7[4:])
    np.testing.assert_array_equal(po.ifftshift(s7), x7)
    np.testing.assert_array_equal(po.ifftshift(s), x)


# ---------------------------------------------------------------- SO(3)
def _wigner_d_explicit(l, m1, m2, beta):
    """Wigner's explicit sum (Sakurai convention), L2-normalised by
    sqrt((2l+1)/2)."""
    f = math.factorial
    pre = math.sqrt(f(l + m1) * f(l - m1) * f(l + m2) * f(l - m2))
    s = 0.0
    c2, s2 = np.cos(beta / 2), np.sin(beta / 2)
    for k in range(max(0, m2 - m1), min(l + m2, l - m1) + 1):
        den = f(l + m2 - k) * f(k) * f(m1 - m2 + k) * f(l - m1 - k)
        s = s + (-1) ** (m1 - m2 + k) / den * c2 ** (2 * l + m2 - m1 - 2 * k) * s2 ** (m1 - m2 + 2 * k)
    return math.sqrt((2 * l + 1) / 2) * pre * s


def test_wigner_d_against_explicit_formula(po):
    bw = 7
    beta = np.pi * (2 * np.arange(2 * bw) + 1) / (4 * bw)
    for m1 in range(-(bw - 1), bw):
        for m2 in range(-(bw - 1), bw):
            d = po.wigner_d(bw, m1, m2)
            l0 = max(abs(m1), abs(m2))
            for l in range(l0, bw):
                np.testing.assert_allclose(d[l - l0], _wigner_d_explicit(l, m1, m2, beta),
                                           rtol=1e-9, atol=1e-12)


def test_so3_correlation_is_real_for_real_signals(po):
    bw = 8
    rng = np.random.default_rng(3)
    n = 2 * bw
    f = po.fst(bw, rng.normal(size=(n, n)))
    h = po.fst(bw, rng.normal(size=(n, n)))
    mag, cp = po.so3_correlate(bw, f, h, want_complex=True)
    assert np.abs(cp.imag).max() < 1e-10 * np.abs(cp.real).max()


def _rot_angle(Ra, Rb):
    c = (np.trace(Ra.T @ Rb) - 1) / 2
    return math.degrees(math.acos(max(-1.0, min(1.0, c))))


@pytest.mark.slow
@pytest.mark.parametrize("seed", [0, 1])
def test_pure_rotation_is_recovered(po, seed):
    """End-to-end pin of the SO(3) conventions (combine signs, Wigner-d,
    alpha/gamma layout, ZYZ decode, quaternion, XYZ Euler with the reference's
    Rz): a cloud and its exact rotation register to that rotation within the
    B=32 bin size (5.6 deg) and zero translation."""
    from phaser_amd import synth
    s, _, R, _ = synth.pair(seed=seed, **synth.CONFIG_B32)
    tg = s.copy()
    tg[:, :3] = (s[:, :3].astype(np.float64) @ R.T).astype(np.float32)
    r = po.register(tg, s, bw=32, n_voxels=64)
    T = r["T_rot"].astype(np.float64)[:3, :3]
    assert _rot_angle(T, R) < 5.0
    assert np.all(r["translation"] == 0)


def test_zyz_from_index_closed_form(po):
    bw = 8
    n = 2 * bw
    for (b, a, g) in [(0, 0, 0), (3, 5, 7), (n - 1, n - 1, n - 1)]:
        idx = b * n * n + a * n + g
        z = po.zyz_from_index(idx, bw)
        np.testing.assert_allclose(z, [np.pi * a / bw, np.pi * (2 * b + 1) / (4 * bw), np.pi * g / bw])


def test_quaternion_to_euler_transform(po):
    """ZYZ -> quaternion -> XYZ Euler -> Tx Ty Tz' reproduces Rz(a)Ry(b)Rz(c)."""
    rng = np.random.default_rng(1)
    for _ in range(20):
        zyz = rng.uniform([0, 0, 0], [2 * np.pi, np.pi, 2 * np.pi])
        q = po.quat_from_zyz(zyz)
        e = po.xyz_from_quat(q)
        T = po.transform_xyz(*e.astype(np.float32))

        def Rz(a):
            return np.array([[np.cos(a), -np.sin(a), 0], [np.sin(a), np.cos(a), 0], [0, 0, 1]])

        def Ry(a):
            return np.array([[np.cos(a), 0, np.sin(a)], [0, 1, 0], [-np.sin(a), 0, np.cos(a)]])
        Rref = Rz(zyz[0]) @ Ry(zyz[1]) @ Rz(zyz[2])
        np.testing.assert_allclose(T[:3, :3], Rref, atol=2e-6)


# ---------------------------------------------------------------- voxels
def test_sub2ind_kats(po):
    # signal-utils-test.cc:210-215
    assert po.sub2ind(100, 100, 100, 201, 201) == 4060300
    assert po.sub2ind(120, 120, 120, 201, 201) == 4872360
    assert po.sub2ind(80, 80, 80, 201, 201) == 3248240


def test_linspaced_and_histc(po):
    E = po.linspaced(128, -50.0, 50.0)
    np.testing.assert_allclose(E, np.linspace(-50, 50, 128), atol=1e-5)
    assert E[-1] == 50.0 and E[0] == -50.0
    assert po.lib().po_histc(50.0, E, 128) == 127
    assert po.lib().po_histc(-50.0, E, 128) == 0
    assert po.lib().po_histc(float(E[5]), E, 128) == 5
    assert po.lib().po_histc(float(np.nextafter(E[5], -np.inf)), E, 128) == 4
    assert po.lib().po_histc(50.0001, E, 128) == -1


def test_voxelize_mean_per_voxel(po):
    N = 4
    pts = np.array([[-49, -49, -49, 0.2], [-48, -48, -48, 0.4], [49, 10, -20, 1.0]], np.float32)
    ch, hist = po.voxelize(pts, N, hist=True)
    E = po.linspaced(N, -50, 50)
    bx = [po.lib().po_histc(float(v), E, N) for v in (49.0, 10.0, -20.0)]
    lin = bx[0] * N + bx[1] + N * N * bx[2]
    assert hist[0] == 2 and hist[lin] == 1
    assert abs(ch[0, 0] - 0.3) < 1e-7
    r0 = np.sqrt(np.float32(3 * 49 * 49))
    assert abs(ch[1, 0] - (np.float64(np.float32(np.sqrt(np.float32(2401) * 3))) +
                            np.float64(np.sqrt(np.float32(3 * 48 * 48)))) / 2) < 1e-4
    assert r0 > 0 and np.all(ch[2] == 0) and np.all(ch[3] == 0)


# ---------------------------------------------------------------- spatial
@pytest.mark.parametrize("N", [3, 4, 5, 8])
def test_r2c_c2r_against_numpy(po, N):
    rng = np.random.default_rng(N)
    x = rng.normal(size=(N, N, N))
    F = po.r2c3(x)
    h = N // 2 + 1
    np.testing.assert_allclose(F[: N * N * h].reshape(N, N, h), np.fft.rfftn(x), atol=1e-10)
    assert np.all(F[N * N * h:] == 0)
    back = po.c2r3(F, N)
    np.testing.assert_allclose(back, x.reshape(-1) * N ** 3, atol=1e-9)


@pytest.mark.parametrize("N", [3, 4, 5])
def test_spatial_correlation_is_circular_cross_correlation(po, N):
    rng = np.random.default_rng(10 + N)
    f = rng.normal(size=(N, N, N))
    g = rng.normal(size=(N, N, N))
    c = po.spatial_correlate(f.reshape(-1), g.reshape(-1), N).reshape(N, N, N)
    brute = np.zeros((N, N, N))
    for s in np.ndindex(N, N, N):
        brute[s] = N ** 3 * np.sum(np.roll(f, [-k for k in s], axis=(0, 1, 2)) * g)
    np.testing.assert_allclose(c, brute, atol=1e-9)


def test_spatial_correlation_end_to_end_kat(po):
    # spatial-correlation-test.cc:59-79
    f = np.zeros(27)
    g = np.zeros(27)
    f[0] = 1
    g[1] = 1
    c = po.spatial_correlate(f, g, 3)
    xyz = po.ind2sub(int(np.argmax(c)), 3, 3)
    assert list(xyz) == [0, 2, 0]


def test_padded_spatial_correlation_kat(po):
    # spatial-correlation-test.cc:103-135 (argmax index + 1 is the test's own)
    f = np.zeros(27)
    g = np.zeros(27)
    f[0] = 1.0
    g[1] = 1.0
    c = po.spatial_correlate(f, g, 3, pad=2)
    idx = int(np.argmax(c)) + 1
    xyz = po.ind2sub(idx, 7, 7)
    x = po.translation_from_index(xyz[0], 7, 0, 3)
    y = po.translation_from_index(xyz[1], 7, 0, 3)
    z = po.translation_from_index(xyz[2], 7, 0, 3)
    assert x == 0 and z == 0
    assert abs(y - (-1.0)) < 0.2


def test_translation_from_index_closed_form(po):
    N = 128
    assert po.translation_from_index(0, N) == 0.0
    assert po.translation_from_index(64, N) == 64 * 101 / 128
    assert po.translation_from_index(65, N) == (65 - 128) * 101 / 128


# ---------------------------------------------------------------- sampling
def test_kdtree_matches_brute_force(po):
    rng = np.random.default_rng(5)
    pts = rng.normal(size=(3000, 4)).astype(np.float32)
    proj, _ = po.project(pts)
    q = po.sampler_grid(8)
    np.testing.assert_array_equal(po.nn(proj, q, kdtree=True), po.nn(proj, q, kdtree=False))
    # duplicated points: ties resolve to the lowest index on both routes
    dup = np.concatenate([proj, proj])
    a = po.nn(dup, q, kdtree=True)
    np.testing.assert_array_equal(a, po.nn(dup, q, kdtree=False))
    assert np.all(a < 3000)


def test_projection_and_grid_point_selection(po):
    bw = 4
    q = po.sampler_grid(bw)
    assert np.allclose(np.linalg.norm(q, axis=1), 1, atol=1e-6)
    pts = np.concatenate([q * 7.0, np.zeros((len(q), 1))], axis=1).astype(np.float32)
    pts[:, 3] = np.arange(len(q)) / len(q)
    proj, rng = po.project(pts)
    assert np.allclose(rng, 7.0, atol=1e-5)
    nn = po.nn(proj, q)
    np.testing.assert_array_equal(nn, np.arange(len(q)))


def test_register_is_reentrant_across_threads(po):
    """The CPU baseline runs oracle registrations on several host threads at
    once (ctypes releases the GIL): results equal the sequential ones."""
    import concurrent.futures as cf
    from phaser_amd import synth
    pairs = [synth.pair(seed=s, rings=16, azimuths=256, noise_sigma=0.02) for s in range(8)]
    seq = [po.register(t, s, bw=8, n_voxels=16) for s, t, _, _ in pairs]
    with cf.ThreadPoolExecutor(max_workers=8) as ex:
        par = list(ex.map(lambda p: po.register(p[1], p[0], bw=8, n_voxels=16), pairs))
    for a, b in zip(seq, par):
        assert (a["rot_index"], a["trans_index"]) == (b["rot_index"], b["trans_index"])
        assert a["rot_value"] == b["rot_value"] and a["trans_value"] == b["trans_value"]
