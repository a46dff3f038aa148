"""ctypes wrapper around the FP64 oracle (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module; it is the checker, never the product path.  The C functions are
described in phaser_oracle.h; the peak-window fits below (peak_window ...
refine_translation, numpy FP64) are pinned by tests/test_peak_fit.py: the
known answers of the reference's bingham-test.cc and gaussian-test.cc, the
window rules case by case, and closed forms.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# PHZ_ORACLE_LIB names another build of phaser_oracle.c (the mutation tests
# load deliberately broken builds through it, to show the pins fail them)
_LIB_PATH = os.environ.get("PHZ_ORACLE_LIB") or os.path.join(HERE, "libphaser_oracle.so")
_lib = None

dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
fp = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
u8p = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")


class Config(C.Structure):
    _fields_ = [("bw", C.c_int), ("n_voxels", C.c_int), ("lower", C.c_int),
                ("upper", C.c_int), ("lp_lower", C.c_int), ("lp_upper", C.c_int),
                ("use_kdtree", C.c_int), ("sph_pad", C.c_int), ("method", C.c_int),
                ("skip_rot_axes", C.c_int), ("spat_pad", C.c_int)]


class Result(C.Structure):
    _fields_ = [("rot_index", C.c_uint32), ("trans_index", C.c_uint32),
                ("rot_value", C.c_double), ("trans_value", C.c_double),
                ("zyz", C.c_double * 3), ("quat", C.c_double * 4),
                ("euler_xyz", C.c_double * 3), ("T_rot", C.c_float * 16),
                ("translation", C.c_double * 3), ("t_sample_ms", C.c_double),
                ("t_rot_ms", C.c_double), ("t_trans_ms", C.c_double)]


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            from oracle import build as _b  # noqa: F401
            _b.build()
        L = C.CDLL(_LIB_PATH)
        sig = {
            "po_fft": (None, [dp, C.c_int, C.c_int64, C.c_int]),
            "po_sampler_grid": (None, [C.c_int, fp]),
            "po_project": (None, [fp, C.c_int64, fp, dp]),
            "po_nn_brute": (None, [fp, C.c_int64, fp, C.c_int, C.c_float, i32p]),
            "po_nn_kdtree": (None, [fp, C.c_int64, fp, C.c_int, C.c_float, i32p]),
            "po_squared_voxel_size": (C.c_float, [C.c_int, C.c_int, C.c_int]),
            "po_sample_cloud": (None, [C.c_int, fp, C.c_int64, C.c_float, C.c_int, dp]),
            "po_sample_cloud_ex": (None, [C.c_int, fp, C.c_void_p, C.c_int64, C.c_float, C.c_int,
                                          dp]),
            "po_seanindex": (C.c_int, [C.c_int, C.c_int, C.c_int]),
            "po_makeweights": (None, [C.c_int, dp]),
            "po_legendre": (None, [C.c_int, C.c_int, dp]),
            "po_fst": (None, [C.c_int, dp, dp, dp]),
            "po_pyramid_struct": (None, [C.c_uint32, C.c_int, C.c_float, u32p, u32p, u32p]),
            "po_fuse_channels": (None, [dp, C.c_int, C.c_uint32, C.c_int, C.c_float, dp]),
            "po_fusion_modes": (None, [C.c_uint32, C.c_int, C.c_float, u8p]),
            "po_fftshift": (None, [dp, C.c_size_t]),
            "po_ifftshift": (None, [dp, C.c_size_t]),
            "po_fftshift_real": (None, [dp, C.c_size_t]),
            "po_ifftshift_real": (None, [dp, C.c_size_t]),
            "po_wigner_d": (None, [C.c_int, C.c_int, C.c_int, dp]),
            "po_so3_correlate": (None, [C.c_int, dp, dp, dp, dp, C.c_void_p, dp]),
            "po_so3_synthesis": (None, [C.c_int, dp, dp, dp, dp, dp]),
            "po_spherical_correlate_laplace": (None, [C.c_int, C.c_int, dp, dp, dp, C.c_void_p, C.c_void_p]),
            "po_spherical_correlate_blend": (None, [C.c_int, dp, dp, dp, dp, dp]),
            "po_linspaced": (None, [C.c_int, C.c_float, C.c_float, fp]),
            "po_histc": (C.c_int, [C.c_double, fp, C.c_int]),
            "po_voxelize": (None, [fp, C.c_int64, C.c_int, C.c_float, C.c_float, dp, C.c_void_p]),
            "po_voxelize_ex": (None, [fp, C.c_void_p, C.c_int64, C.c_int, C.c_float, C.c_float, dp,
                                      C.c_void_p]),
            "po_r2c3": (None, [C.c_int, dp, dp]),
            "po_c2r3": (None, [C.c_int, dp, dp]),
            "po_spatial_correlate": (None, [C.c_int, C.c_int, dp, dp, dp]),
            "po_spatial_correlate_laplace": (None, [C.c_int, C.c_int, dp, dp, dp]),
            "po_padded_size": (C.c_uint32, [C.c_int, C.c_int]),
            "po_spatial_pad_inverse": (None, [C.c_int, C.c_int, C.c_int, dp, dp]),
            "po_spatial_correlate_laplace_pad": (None, [C.c_int, C.c_int, C.c_int, dp, dp, dp]),
            "po_argmax": (C.c_uint32, [dp, C.c_size_t]),
            "po_zscore_peaks": (C.c_size_t, [dp, C.c_size_t, C.c_double, C.c_uint32, C.c_double,
                                             C.c_double, C.c_double, u32p, C.c_size_t]),
            "po_zyz_from_index": (None, [C.c_uint32, C.c_int, dp]),
            "po_quat_from_zyz": (None, [dp, dp]),
            "po_xyz_from_quat": (None, [dp, dp]),
            "po_transform_xyz": (None, [C.c_float, C.c_float, C.c_float, fp]),
            "po_apply_transform": (None, [fp, fp, C.c_int64, fp]),
            "po_translation_from_index": (C.c_double, [C.c_double, C.c_uint32, C.c_int, C.c_int]),
            "po_ind2sub": (None, [C.c_uint32, C.c_uint32, C.c_uint32, u32p]),
            "po_sub2ind": (C.c_uint32, [C.c_uint32] * 5),
            "po_register": (None, [C.POINTER(Config), fp, C.c_int64, fp, C.c_int64,
                                   C.POINTER(Result), C.c_void_p, C.c_void_p, C.c_void_p]),
            "po_register_ex": (None, [C.POINTER(Config), fp, C.c_void_p, C.c_int64, fp,
                                      C.c_void_p, C.c_int64, C.POINTER(Result), C.c_void_p,
                                      C.c_void_p, C.c_void_p]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


# ---- thin numpy-level helpers -------------------------------------------
def fft(x, sign=-1):
    a = _c(np.asarray(x, np.complex128).view(np.float64), np.float64).copy()
    lib().po_fft(a, len(a) // 2, 1, sign)
    return a.view(np.complex128)


def sampler_grid(bw):
    out = np.zeros((4 * bw * bw, 3), np.float32)
    lib().po_sampler_grid(bw, out)
    return out


def project(pts):
    pts = _c(pts, np.float32)
    P = pts.shape[0]
    proj = np.zeros((P, 3), np.float32)
    rng = np.zeros(P, np.float64)
    lib().po_project(pts, P, proj, rng)
    return proj, rng


def nn(proj, queries, max_sq=1e30, kdtree=False):
    proj = _c(proj, np.float32)
    queries = _c(queries, np.float32)
    out = np.zeros(queries.shape[0], np.int32)
    f = lib().po_nn_kdtree if kdtree else lib().po_nn_brute
    f(proj, proj.shape[0], queries, queries.shape[0], max_sq, out)
    return out


def squared_voxel_size(lp_lower=0, lp_upper=100000, n_voxels=100):
    return lib().po_squared_voxel_size(lp_lower, lp_upper, n_voxels)


def _extra(extra, P):
    """(reflectivity, ambient) per point as a contiguous (P, 2) float32 array,
    or None."""
    if extra is None:
        return None
    e = _c(extra, np.float32)
    assert e.shape == (P, 2), e.shape
    return e


def sample_cloud(bw, pts, max_sq=None, kdtree=True, extra=None):
    """extra: (P, 2) reflectivity and ambient per point, or None (zero
    channels, as for a PointXYZI cloud)."""
    pts = _c(pts, np.float32)
    if max_sq is None:
        max_sq = squared_voxel_size()
    out = np.zeros((4, 4 * bw * bw), np.float64)
    e = _extra(extra, pts.shape[0])
    lib().po_sample_cloud_ex(bw, pts, _ptr(e), pts.shape[0], max_sq, int(kdtree), out)
    return out


def seanindex(m, l, bw):
    return lib().po_seanindex(m, l, bw)


def makeweights(bw):
    w = np.zeros(4 * bw, np.float64)
    lib().po_makeweights(bw, w)
    return w


def legendre(bw, m):
    out = np.zeros((bw - m, 2 * bw), np.float64)
    lib().po_legendre(bw, m, out)
    return out


def fst(bw, data):
    data = _c(data, np.float64).reshape(-1)
    re = np.zeros(bw * bw)
    im = np.zeros(bw * bw)
    lib().po_fst(bw, data, re, im)
    return re + 1j * im


def pyramid_struct(n, levels, div=4.0):
    a = np.zeros(levels, np.uint32)
    b = np.zeros(levels, np.uint32)
    c = np.zeros(levels, np.uint32)
    lib().po_pyramid_struct(n, levels, div, a, b, c)
    return a, b, c


def fuse_channels(channels, levels=5, div=4.0):
    ch = _c(np.asarray(channels, np.complex128), np.complex128)
    Cn, n = ch.shape
    out = np.zeros(n, np.complex128)
    lib().po_fuse_channels(ch.view(np.float64), Cn, n, levels, div, out.view(np.float64))
    return out


def fusion_modes(n, levels=5, div=4.0):
    out = np.zeros(n, np.uint8)
    lib().po_fusion_modes(n, levels, div, out)
    return out


def fftshift(x):
    a = _c(np.asarray(x, np.complex128), np.complex128).copy()
    lib().po_fftshift(a.view(np.float64), len(a))
    return a


def ifftshift(x):
    a = _c(np.asarray(x, np.complex128), np.complex128).copy()
    lib().po_ifftshift(a.view(np.float64), len(a))
    return a


def wigner_d(bw, m1, m2):
    l0 = max(abs(m1), abs(m2))
    out = np.zeros((bw - l0, 2 * bw), np.float64)
    lib().po_wigner_d(bw, m1, m2, out)
    return out


def so3_correlate(bw, f, h, want_complex=False):
    f = _c(f, np.complex128)
    h = _c(h, np.complex128)
    n = 2 * bw
    mag = np.zeros(n ** 3)
    cp = np.zeros(n ** 3, np.complex128) if want_complex else None
    lib().po_so3_correlate(bw, _c(f.real, np.float64), _c(f.imag, np.float64),
                           _c(h.real, np.float64), _c(h.imag, np.float64),
                           _ptr(cp.view(np.float64)) if cp is not None else None, mag)
    return (mag, cp) if want_complex else mag


def so3_synthesis(bw, f, h):
    n = 2 * bw
    S = np.zeros((n, n, n), np.complex128)
    f = _c(f, np.complex128)
    h = _c(h, np.complex128)
    lib().po_so3_synthesis(bw, _c(f.real, np.float64), _c(f.imag, np.float64),
                           _c(h.real, np.float64), _c(h.imag, np.float64),
                           S.view(np.float64).reshape(-1))
    return S


def spherical_correlate_blend(bw, f, h):
    """SphericalCorrelation::correlateSignals (spherical-correlation.cc:54-80,
    103-112) on sampled grids f, h (4 x (2bw)^2, channel order range,
    intensity, reflectivity, ambient): |C|^2 on the (2bw)^3 grid of the
    channel 0.6 intensity + 0.4 range."""
    n = 2 * bw
    f = np.ascontiguousarray(f, np.float64).reshape(4, n * n)
    h = np.ascontiguousarray(h, np.float64).reshape(4, n * n)
    mag = np.zeros(n ** 3)
    lib().po_spherical_correlate_blend(bw, f[0].copy(), f[1].copy(), h[0].copy(), h[1].copy(),
                                       mag)
    return mag


def spherical_correlate_laplace(bw, f, h, fused=False):
    f = _c(f, np.float64)
    h = _c(h, np.float64)
    Cn = f.shape[0]
    mag = np.zeros((2 * bw) ** 3)
    ff = np.zeros(bw * bw, np.complex128)
    hf = np.zeros(bw * bw, np.complex128)
    lib().po_spherical_correlate_laplace(bw, Cn, f, h, mag, _ptr(ff.view(np.float64)),
                                         _ptr(hf.view(np.float64)))
    return (mag, ff, hf) if fused else mag


def linspaced(N, lo, hi):
    E = np.zeros(N, np.float32)
    lib().po_linspaced(N, lo, hi, E)
    return E


def voxelize(pts, N, lo=-50.0, hi=50.0, hist=False, extra=None):
    pts = _c(pts, np.float32)
    ch = np.zeros((4, N ** 3))
    hs = np.zeros(N ** 3) if hist else None
    e = _extra(extra, pts.shape[0])
    lib().po_voxelize_ex(pts, _ptr(e), pts.shape[0], N, lo, hi, ch, _ptr(hs))
    return (ch, hs) if hist else ch


def r2c3(x):
    x = _c(x, np.float64)
    N = round(len(x.reshape(-1)) ** (1 / 3))
    out = np.zeros(N ** 3, np.complex128)
    lib().po_r2c3(N, x.reshape(-1), out.view(np.float64))
    return out


def c2r3(X, N):
    X = _c(X, np.complex128)
    out = np.zeros(N ** 3)
    lib().po_c2r3(N, X.view(np.float64), out)
    return out


def spatial_correlate(f, g, N, pad=0):
    out = np.zeros((N + 2 * pad) ** 3)
    lib().po_spatial_correlate(N, pad, _c(f, np.float64), _c(g, np.float64), out)
    return out


def spatial_correlate_laplace(f, g, N, pad=0):
    f = _c(f, np.float64)
    g = _c(g, np.float64)
    out = np.zeros((N + 2 * pad) ** 3)
    lib().po_spatial_correlate_laplace_pad(N, pad, f.shape[0], f, g, out)
    return out


def spherical_correlate_low_pass(bw, f_grid, h_grid, lower, upper):
    """SphericalCorrelationLowPass::correlateSampledSignals (spherical-
    correlation-low-pass.cc:14-59): the S^2 transforms of one grid per
    signal, FFTShifted over B^2 (both parts alike), zero outside
    [min(B^2, lower), min(B^2, upper)), IFFTShifted back, then the SO(3)
    correlation |C|^2 (the combine and inverse of po_so3_correlate)."""
    if lower < 0 or upper <= 0:
        raise PeakWindowError("low-pass bounds fail the reference's CHECKs (:26-27)")
    n = 2 * bw
    full = bw * bw
    lo, hi = min(full, lower), min(full, upper)
    coefs = []
    for grid in (f_grid, h_grid):
        c = fftshift(fst(bw, np.asarray(grid, np.float64).reshape(n, n)))
        c[:lo] = 0
        c[hi:] = 0
        coefs.append(ifftshift(c))
    return so3_correlate(bw, coefs[0], coefs[1])


def spatial_correlate_low_pass(f, g, N, lower, upper, pad=0):
    """SpatialCorrelationLowPass::correlateSignals (spatial-correlation-low-
    pass.cc:21-94) on its first call (a fresh object: zero product buffer and
    zero spectrum tails): the r2c spectra of the one channel, FFTShifted over
    N^3; products only at the shifted indices Sub2Ind(i, j, k, N, N) = i N +
    j + N^2 k, i, j, k in [lower, min(upper, N)]; the padded inverse with
    IFFTShift (po_spatial_pad_inverse).  An index >= N^3 fails the CHECK_LT of
    complexMulSeqUsingIndices (spatial-correlation.cc:153), an empty cube its
    CHECK (:145)."""
    hi = min(upper, N)
    if lower > hi:
        raise PeakWindowError("empty low-pass cube (spatial-correlation.cc:145)")
    if hi >= N:
        raise PeakWindowError("low-pass cube reaches N^3 (spatial-correlation.cc:153)")
    F = fftshift(r2c3(np.asarray(f, np.float64).reshape(-1)))
    G = fftshift(r2c3(np.asarray(g, np.float64).reshape(-1)))
    r = np.arange(lower, hi + 1)
    i, j, k = np.meshgrid(r, r, r, indexing="ij")
    idx = (i * N + j + N * N * k).reshape(-1)
    P = np.zeros(N ** 3, np.complex128)
    P[idx] = F[idx] * np.conj(G[idx])
    return spatial_pad_inverse(P, N, pad, True)


def padded_size(N, pad):
    return int(lib().po_padded_size(N, pad))


def spatial_pad_inverse(prod, N, pad, shifted):
    """prod: N^3 complex products in their own order -> (N + 2 pad)^3 grid."""
    p = _c(np.asarray(prod, np.complex128).view(np.float64), np.float64)
    out = np.zeros((N + 2 * pad) ** 3)
    lib().po_spatial_pad_inverse(N, pad, int(shifted), p, out)
    return out


def zyz_from_index(idx, bw):
    z = np.zeros(3)
    lib().po_zyz_from_index(idx, bw, z)
    return z


def quat_from_zyz(zyz):
    q = np.zeros(4)
    lib().po_quat_from_zyz(_c(zyz, np.float64), q)
    return q


def xyz_from_quat(q):
    e = np.zeros(3)
    lib().po_xyz_from_quat(_c(q, np.float64), e)
    return e


def transform_xyz(a, b, c):
    T = np.zeros(16, np.float32)
    lib().po_transform_xyz(a, b, c, T)
    return T.reshape(4, 4)


def apply_transform(T, pts):
    pts = _c(pts, np.float32)
    out = np.zeros_like(pts)
    lib().po_apply_transform(_c(T, np.float32).reshape(-1), pts, pts.shape[0], out)
    return out


def translation_from_index(idx, N, lo=-50, hi=50):
    return lib().po_translation_from_index(float(idx), N, lo, hi)


class PeakWindowError(ValueError):
    """A case where the reference fails a CHECK (the process would abort)."""


def peak_window(idx, n_corr, k, kind):
    """Linear-index neighbours [start, end] of the peak, as
    BinghamPeakBasedEval::calculateStartEndNeighbor
    (bingham-peak-based-eval.cc:47-58) and its Gaussian counterpart
    (gaussian-peak-based-eval.cc:26-37): no lower neighbours at all when
    k > idx; for the Bingham fit end = idx when idx + k > n_corr, and
    end == n_corr fails the reference's CHECK_LT(end, n) (:69); for the
    Gaussian end = idx when idx + k >= n_corr.  The mixture evaluations:
    "bmm" as "bingham" (bmm-peak-based-eval.cc:46-56, CHECK_LT :68), "gmm"
    with end = idx when idx + k > n_corr and CHECK_LT(end, n) (gmm-peak-
    based-eval.cc:58-68, :81)."""
    start = idx if k > idx else idx - k
    if kind in ("bingham", "bmm", "gmm"):
        end = idx if k + idx > n_corr else idx + k
        if end >= n_corr:
            raise PeakWindowError("peak window reaches past the grid")
    else:
        end = idx if k + idx >= n_corr else idx + k
    return start, end


def bingham_moment(samples, weights):
    """The moment matrix C = S diag(w) S^T, symmetrised (bingham.cc:281-296)."""
    S = np.asarray(samples, np.float64)
    w = np.asarray(weights, np.float64)
    Cm = S @ np.diag(w) @ S.T
    return 0.5 * (Cm + Cm.T)


def bingham_mode(samples, weights, ref=None):
    """Bingham::fit -> fitToMoment -> mode (bingham.cc:281-296, 239-279,
    110-113): the eigenvectors of the moment matrix are ordered by their
    maximum-likelihood concentrations, which order like the eigenvalues, so
    the mode is the eigenvector of the largest eigenvalue.  Its sign is free
    (the pose depends on q q^T only): here the one with m . ref >= 0, or
    m[0] >= 0 without ref."""
    _, vecs = np.linalg.eigh(bingham_moment(samples, weights))
    m = vecs[:, -1]
    flip = (m @ np.asarray(ref, np.float64) < 0) if ref is not None else m[0] < 0
    return -m if flip else m


def rotation_window(bw, corr, idx, k, kind="bingham"):
    """BinghamPeakBasedEval::retrievePeakNeighbors (bingham-peak-based-eval.cc:
    60-90) on the window of peak_window: the quaternions (4 x m) of the
    window's indices (rotation-utils.cc:101-107,190-207) and their weights,
    the correlation values, normalised to sum 1 when m > 1 (CHECK_GT of the
    sum, :88); also the peak's own quaternion.  kind "bmm"
    (BmmPeakBasedEval::retrievePeakNeighbors, bmm-peak-based-eval.cc:58-86)
    normalises and checks one sample too."""
    start, end = peak_window(idx, len(corr), k, kind)
    peak = quat_from_zyz(zyz_from_index(idx, bw))
    samples = np.stack([quat_from_zyz(zyz_from_index(i, bw)) for i in range(start, end + 1)],
                       axis=1)
    w = np.asarray(corr[start:end + 1], np.float64)
    if end > start or kind == "bmm":
        if not w.sum() > 0:
            raise PeakWindowError("weights do not sum to a positive value")
        w = w / w.sum()
    return samples, w, peak


def refine_rotation(bw, corr, idx, k, kind="bingham"):
    """BinghamPeakBasedEval::fitRotationalBinghamDistribution with k
    neighbours (bingham-peak-based-eval.cc:22-45): returns the mode quaternion
    (w, x, y, z), signed to agree with the peak's quaternion, and omega, the
    moment matrix's eigenvalues over their sum in ascending order
    (fitToMoment, bingham.cc:239-246).  One sample: the mode of w q q^T is q
    itself and omega = (0, 0, 0, 1)."""
    samples, w, peak = rotation_window(bw, corr, idx, k, kind)
    if samples.shape[1] == 1:
        return peak, np.array([0.0, 0.0, 0.0, 1.0])
    ev = np.linalg.eigvalsh(bingham_moment(samples, w))
    return bingham_mode(samples, w, ref=peak), ev / ev.sum()


def gaussian_fit(samples, weights):
    """Gaussian(samples, weights) -> setMeanAndCov (gaussian.cc:14-22, 48-56):
    mean = S w; covariance = sum_i w_i (s_i - mean)(s_i - mean)^T, formed
    from sqrt(w_i)-scaled columns, so a negative weight makes it NaN."""
    S = np.asarray(samples, np.float64)
    w = np.asarray(weights, np.float64)
    mu = S @ w
    with np.errstate(invalid="ignore"):
        Z = (S - mu[:, None]) * np.sqrt(w)[None, :]
    return mu, Z @ Z.T


def translation_window(N, corr, idx, k, lo=-50, hi=50, kind="gaussian"):
    """GaussianPeakBasedEval::retrievePeakNeighbors (gaussian-peak-based-eval.cc:
    75-102) on the window of peak_window: the translations (3 x m) of the
    window's indices (signal-utils.cc:11-19, translation-utils.cc:40-52) and
    the correlation values normalised to sum 1 (CHECK_GT of the sum, :100)."""
    start, end = peak_window(idx, len(corr), k, kind)
    t = []
    for i in range(start, end + 1):
        x, y, z = ind2sub(i, N, N)
        t.append([translation_from_index(x, N, lo, hi), translation_from_index(y, N, lo, hi),
                  translation_from_index(z, N, lo, hi)])
    w = np.asarray(corr[start:end + 1], np.float64)
    if not w.sum() > 0:
        raise PeakWindowError("weights do not sum to a positive value")
    return np.asarray(t, np.float64).T, w / w.sum()


def refine_translation(N, corr, idx, k, lo=-50, hi=50, kind="gaussian"):
    """GaussianPeakBasedEval::fitTranslationalNormalDist with k neighbours
    (gaussian-peak-based-eval.cc:39-73): (mean, covariance) of the window.
    kind "gmm": GmmPeakBasedEval's window (gmm-peak-based-eval.cc:58-105);
    its one-signal mixture is that Gaussian (gaussian_mixture)."""
    return gaussian_fit(*translation_window(N, corr, idx, k, lo, hi, kind))


def gaussian_mixture(means, covs, weights):
    """GaussianMixture::calcMixutreParameters (gaussian-mixture.cc:57-78): the
    mean and covariance of the component means under the weights (a
    Gaussian(means, weights), gaussian.cc:14-22,48-56) plus the weighted sum
    of the component covariances."""
    mu, cov_mean = gaussian_fit(means, weights)
    w = np.asarray(weights, np.float64)
    return mu, sum(np.asarray(c, np.float64) * w[i] for i, c in enumerate(covs)) + cov_mean


def refine_translation_gmm(N, corr, signals, k, lo=-50, hi=50):
    """GmmPeakBasedEval::fitTranslationalGmmDistribution (gmm-peak-based-eval.
    cc:27-56): one Gaussian per signal over its window, mixture weights the
    signals' correlation values normalised to sum 1."""
    comps = [refine_translation(N, corr, i, k, lo, hi, "gmm") for i in signals]
    w = np.asarray([corr[i] for i in signals], np.float64)
    w = w / w.sum()
    return gaussian_mixture(np.stack([m for m, _ in comps], axis=1), [c for _, c in comps], w)


ZSCORE_DEFAULTS = dict(lag_percentile=0.05, lag_max=20, threshold=4.33, influence=0.05,
                      filter_threshold=0.405)  # z-score-peak-extraction.cc:12-25


def zscore_peaks(corr, lag_percentile=0.05, lag_max=20, threshold=4.33, influence=0.05,
                 filter_threshold=0.405, max_peaks=1 << 20):
    """ZScorePeakExtraction::extractPeaks (po_zscore_peaks): the signal
    indices in increasing order."""
    c = _c(corr, np.float64)
    out = np.zeros(max_peaks, np.uint32)
    k = lib().po_zscore_peaks(c, c.size, lag_percentile, lag_max, threshold, influence,
                              filter_threshold, out, max_peaks)
    return out[: min(k, max_peaks)].copy()


def voxel_grid_filter(pts, leaf=0.25):
    """VoxelGridCmd (voxel-grid-cmd.cc:11-24): PCL 1.8 VoxelGrid::applyFilter
    (voxel_grid.hpp) with every field averaged, in float32 as PCL: the
    finite points' bounding box, the overflow check (more than INT32_MAX
    voxels: the cloud unchanged), min_b = floor(min * inv), the voxel index
    (int)(floor(p * inv) - (float)min_b) . (1, div_x, div_x div_y), the
    points sorted by index (stably: each voxel's points in input order; PCL's
    std::sort may take them in another order), one centroid per voxel (float
    sums in order, divided by the count), ordered by voxel index.
    pts: (P, 4) x, y, z, intensity."""
    a = np.asarray(pts, np.float32).reshape(-1, 4)
    fin = np.all(np.isfinite(a[:, :3]), axis=1)
    if not fin.any():
        return np.zeros((0, 4), np.float32)
    inv = np.float32(1.0) / np.float32(leaf)
    mn = a[fin, :3].min(axis=0)
    mx = a[fin, :3].max(axis=0)
    d = [int(np.float32((mx[k] - mn[k]) * inv)) + 1 for k in range(3)]
    if d[0] * d[1] * d[2] > 2 ** 31 - 1:
        return a.copy()
    min_b = np.floor(mn * inv).astype(np.int64)
    max_b = np.floor(mx * inv).astype(np.int64)
    div = max_b - min_b + 1
    q = a[fin]
    ijk = (np.floor(q[:, :3] * inv) - min_b.astype(np.float32)).astype(np.int64)
    idx = ijk[:, 0] + ijk[:, 1] * div[0] + ijk[:, 2] * div[0] * div[1]
    order = np.argsort(idx, kind="stable")
    q = q[order]
    keys, start, count = np.unique(idx[order], return_index=True, return_counts=True)
    sums = np.zeros((len(keys), 4), np.float32)
    for k in range(int(count.max())):  # sequential float32 sums per voxel
        live = count > k
        sums[live] += q[start[live] + k]
    return (sums / count.astype(np.float32)[:, None]).astype(np.float32)


def pass_through_filter(pts, field=2, lower=-5.0, upper=3.0, negative=True):
    """PassThroughGndFilterCmd (pass-through-gnd-filter-cmd.cc:14-26): PCL 1.8
    PassThrough::applyFilterIndices (passthrough.hpp): non-finite points and
    non-finite field values dropped; kept, in input order, the points whose
    field (0 x, 1 y, 2 z, 3 intensity) lies in [lower, upper] (float), or
    outside it with negative (the command: z, -5, 3, negative)."""
    a = np.asarray(pts, np.float32).reshape(-1, 4)
    v = a[:, field]
    lo, hi = np.float32(lower), np.float32(upper)
    ok = np.all(np.isfinite(a[:, :3]), axis=1) & np.isfinite(v)
    inside = (v >= lo) & (v <= hi)
    keep = ok & (~inside if negative else inside)
    return a[keep].copy()


def ind2sub(idx, rows, cols):
    out = np.zeros(3, np.uint32)
    lib().po_ind2sub(idx, rows, cols, out)
    return out


def sub2ind(i, j, k, rows, cols):
    return lib().po_sub2ind(i, j, k, rows, cols)


def register(target, source, bw=32, n_voxels=64, lower=-50, upper=50,
             kdtree=True, want_grids=False, want_reg=False, target_extra=None,
             source_extra=None, sph_pad=0, method=0, skip_rot_axes=0, lp_lower=0,
             lp_upper=100000, spat_pad=0):
    """target_extra / source_extra: (P, 2) reflectivity and ambient per
    point, or None (zero channels).  sph_pad: spherical zero padding (the
    SO(3) grid at bw + sph_pad).  method 1: SphRegistration's single blended
    channel (no padding); skip_rot_axes: its refine_rot_x/y/z masks.
    lp_lower / lp_upper: the spatial low-pass bounds (they set the S^2
    nearest-neighbour acceptance radius).  spat_pad: spatial zero padding
    (translation grid (n_voxels + 2 spat_pad)^3)."""
    cfg = Config(bw, n_voxels, lower, upper, int(lp_lower), int(lp_upper), int(kdtree),
                 int(sph_pad), int(method), int(skip_rot_axes), int(spat_pad))
    if method == 1:
        sph_pad = 0
    t = _c(target, np.float32)
    s = _c(source, np.float32)
    te, se = _extra(target_extra, t.shape[0]), _extra(source_extra, s.shape[0])
    res = Result()
    rot = np.zeros((2 * (bw + sph_pad)) ** 3) if want_grids else None
    tr = np.zeros((n_voxels + 2 * spat_pad) ** 3) if want_grids else None
    reg = np.zeros_like(s) if want_reg else None
    lib().po_register_ex(C.byref(cfg), t, _ptr(te), t.shape[0], s, _ptr(se), s.shape[0],
                         C.byref(res), _ptr(reg), _ptr(rot), _ptr(tr))
    out = {
        "rot_index": res.rot_index, "trans_index": res.trans_index,
        "rot_value": res.rot_value, "trans_value": res.trans_value,
        "zyz": np.array(res.zyz[:]), "quat": np.array(res.quat[:]),
        "euler_xyz": np.array(res.euler_xyz[:]),
        "T_rot": np.array(res.T_rot[:], np.float32).reshape(4, 4),
        "translation": np.array(res.translation[:]),
        "t_ms": (res.t_sample_ms, res.t_rot_ms, res.t_trans_ms),
    }
    if want_grids:
        out["rot_corr"] = rot
        out["trans_corr"] = tr
    if want_reg:
        out["registered"] = reg
    return out
