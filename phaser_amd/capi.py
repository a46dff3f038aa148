"""ctypes binding of the C ABI in include/phaser_amd.h.

Loads the in-tree lib/libphaser_amd.so and fails loudly when it is missing:
there is no CPU fallback on the product path.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libphaser_amd.so")
_lib = None


class Config(C.Structure):
    _fields_ = [("bandwidth", C.c_int32), ("spherical_zero_padding", C.c_int32),
                ("n_voxels", C.c_int32), ("discretize_lower", C.c_int32),
                ("discretize_upper", C.c_int32), ("spatial_zero_padding", C.c_int32),
                ("low_pass_lower", C.c_int32), ("low_pass_upper", C.c_int32),
                ("max_points", C.c_int64), ("max_batch", C.c_int32), ("device", C.c_int32),
                ("precision", C.c_int32), ("bingham_peak_neighbors", C.c_int32),
                ("gaussian_peak_neighbors", C.c_int32), ("channels", C.c_int32),
                ("method", C.c_int32), ("skip_rot_axes", C.c_int32), ("rot_eval", C.c_int32),
                ("pos_eval", C.c_int32), ("bmm_peak_neighbors", C.c_int32),
                ("gmm_peak_neighbors", C.c_int32)]


class Result(C.Structure):
    _fields_ = [("rot_index", C.c_uint32), ("trans_index", C.c_uint32),
                ("rot_value", C.c_float), ("trans_value", C.c_float),
                ("zyz", C.c_double * 3), ("quat", C.c_double * 4),
                ("euler_xyz", C.c_double * 3), ("T_rot", C.c_float * 16),
                ("translation", C.c_double * 3), ("trans_cov", C.c_double * 9),
                ("rot_omega", C.c_double * 4), ("rot_axes", C.c_double * 16),
                ("status", C.c_int32), ("reserved", C.c_int32)]

    def as_dict(self):
        return {
            "rot_index": int(self.rot_index), "trans_index": int(self.trans_index),
            "rot_value": float(self.rot_value), "trans_value": float(self.trans_value),
            "zyz": np.array(self.zyz[:]), "quat": np.array(self.quat[:]),
            "euler_xyz": np.array(self.euler_xyz[:]),
            "T_rot": np.array(self.T_rot[:], np.float32).reshape(4, 4),
            "translation": np.array(self.translation[:]),
            "trans_cov": np.array(self.trans_cov[:]).reshape(3, 3),
            "rot_omega": np.array(self.rot_omega[:]),
            "rot_axes": np.array(self.rot_axes[:]).reshape(4, 4),
            "status": int(self.status),
        }


def _struct_dtype(st):
    """The numpy dtype with the layout of ctypes structure st (scalars and
    one-dimensional arrays of 32-bit ints / floats and doubles)."""
    scalar = {C.c_uint32: np.uint32, C.c_int32: np.int32, C.c_float: np.float32,
              C.c_double: np.float64}
    names, formats, offsets = [], [], []
    for name, ct in st._fields_:
        names.append(name)
        offsets.append(getattr(st, name).offset)
        formats.append((scalar[ct._type_], (ct._length_,)) if hasattr(ct, "_length_")
                       else scalar[ct])
    return np.dtype({"names": names, "formats": formats, "offsets": offsets,
                     "itemsize": C.sizeof(st)})


_RESULT_DT = _struct_dtype(Result)
_RESULT_SHAPES = {"T_rot": (4, 4), "trans_cov": (3, 3), "rot_axes": (4, 4)}


def results_as_dicts(res, n):
    """Result.as_dict() of the first n entries of a ctypes Result array,
    converted column by column (one copy of the block instead of a dozen
    small arrays per result); the same keys, types and values."""
    a = np.frombuffer(res, dtype=_RESULT_DT, count=n).copy()
    cols = {k: a[k].reshape(n, *_RESULT_SHAPES.get(k, a[k].shape[1:]))
            for k in ("zyz", "quat", "euler_xyz", "T_rot", "translation", "trans_cov",
                      "rot_omega", "rot_axes")}
    ri, ti, st = a["rot_index"].tolist(), a["trans_index"].tolist(), a["status"].tolist()
    rv = a["rot_value"].astype(np.float64).tolist()
    tv = a["trans_value"].astype(np.float64).tolist()
    return [{"rot_index": ri[k], "trans_index": ti[k], "rot_value": rv[k], "trans_value": tv[k],
             **{name: col[k] for name, col in cols.items()}, "status": st[k]}
            for k in range(n)]


EXPORTS = {
    "phz_default_config": (None, [C.POINTER(Config)]),
    "phz_create": (C.c_int, [C.POINTER(Config), C.POINTER(C.c_void_p)]),
    "phz_destroy": (None, [C.c_void_p]),
    "phz_last_error": (C.c_char_p, []),
    "phz_get_config": (C.c_int, [C.c_void_p, C.POINTER(Config)]),
    "phz_register": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.c_int64,
                               C.POINTER(Result), C.c_void_p, C.c_void_p, C.c_void_p]),
    "phz_register_device": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.c_int64,
                                      C.POINTER(Result), C.c_void_p, C.c_void_p, C.c_void_p,
                                      C.c_void_p]),
    "phz_register_pairs_device": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_void_p,
                                            C.c_void_p, C.c_void_p, C.POINTER(Result),
                                            C.c_void_p, C.c_void_p, C.c_void_p]),
    "phz_register_pairs": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_void_p,
                                     C.c_void_p, C.c_void_p, C.POINTER(Result)]),
    "phz_register_pairs_channels_device": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_void_p,
                                                     C.c_void_p, C.c_void_p, C.c_void_p,
                                                     C.c_void_p, C.c_void_p, C.POINTER(Result),
                                                     C.c_void_p, C.c_void_p, C.c_void_p]),
    "phz_register_channels": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64,
                                        C.c_void_p, C.c_void_p, C.c_int64, C.POINTER(Result),
                                        C.c_void_p, C.c_void_p, C.c_void_p]),
    "phz_memory_usage": (C.c_int, [C.c_void_p, C.POINTER(C.c_size_t)]),
    "phz_estimate_memory": (C.c_int, [C.POINTER(Config), C.POINTER(C.c_size_t)]),
    "phz_sample_s2": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p]),
    "phz_s2_transform": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]),
    "phz_so3_correlate_coef": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                         C.c_void_p, C.POINTER(C.c_uint32)]),
    "phz_spherical_correlate": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p,
                                          C.c_void_p, C.POINTER(C.c_uint32)]),
    "phz_voxelize": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p]),
    "phz_spatial_correlate": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p,
                                        C.POINTER(C.c_uint32)]),
    "phz_spatial_correlate_single": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p,
                                               C.c_void_p, C.POINTER(C.c_uint32)]),
    "phz_spatial_correlate_single_padded": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_void_p,
                                                      C.c_void_p, C.c_void_p,
                                                      C.POINTER(C.c_uint32)]),
    "phz_spatial_correlate_single_noalloc": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p,
                                                       C.c_void_p, C.c_void_p,
                                                       C.POINTER(C.c_uint32), C.c_void_p]),
    "phz_spatial_correlate_low_pass": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int,
                                                 C.c_void_p, C.c_void_p, C.c_void_p,
                                                 C.POINTER(C.c_uint32)]),
    "phz_spherical_correlate_low_pass": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int,
                                                   C.c_int, C.c_void_p, C.POINTER(C.c_uint32)]),
    "phz_voxel_grid_filter": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int64, C.c_float,
                                        C.POINTER(C.c_int64)]),
    "phz_pass_through_filter": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int64, C.c_int, C.c_float,
                                          C.c_float, C.c_int, C.POINTER(C.c_int64)]),
    "phz_mfma_peak": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.POINTER(C.c_double)]),
    "phz_zscore_default_params": (None, [C.c_void_p]),
    "phz_zscore_peaks": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p,
                                   C.c_int64, C.POINTER(C.c_int64)]),
    "phz_autotune": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p]),
    "phz_get_tuning": (C.c_int, [C.c_void_p, C.c_void_p]),
    "phz_set_tuning": (C.c_int, [C.c_void_p, C.c_void_p]),
    "phz_set_graphs": (C.c_int, [C.c_void_p, C.c_int]),
    "phz_set_profiling": (C.c_int, [C.c_void_p, C.c_int]),
    "phz_get_stage_times": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.POINTER(C.c_int)]),
    "phz_device_malloc": (C.c_int, [C.c_void_p, C.c_size_t, C.POINTER(C.c_void_p)]),
    "phz_device_free": (C.c_int, [C.c_void_p, C.c_void_p]),
    "phz_memcpy": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]),
    "phz_synchronize": (C.c_int, [C.c_void_p]),
    "phz_copy_kernel": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t, C.c_int,
                                  C.POINTER(C.c_float)]),
}


class ZScoreParams(C.Structure):
    _fields_ = [("lag_percentile", C.c_double), ("lag_max", C.c_uint32),
                ("reserved", C.c_uint32), ("threshold", C.c_double), ("influence", C.c_double),
                ("filter_threshold", C.c_double)]


class Tuning(C.Structure):
    _fields_ = [("fwd_x", C.c_int32), ("fwd_z", C.c_int32), ("inv_x", C.c_int32),
                ("plane_fwd", C.c_int32), ("nn_lanes", C.c_int32), ("fused_reg", C.c_int32),
                ("so3_reg", C.c_int32), ("occ_fwd", C.c_int32), ("inv_plane", C.c_int32),
                ("nn_xcd", C.c_int32), ("nn_seed", C.c_int32), ("xz_slab", C.c_int32),
                ("ms_fwd_x", C.c_float), ("ms_fwd_z", C.c_float),
                ("ms_inv_x", C.c_float), ("ms_plane", C.c_float), ("ms_fused", C.c_float),
                ("ms_so3", C.c_float)]

    def as_dict(self):
        return {k: (round(getattr(self, k), 4) if k.startswith("ms") else getattr(self, k))
                for k, _ in self._fields_ if not k.startswith("reserved")}


class StageTime(C.Structure):
    _fields_ = [("name", C.c_char * 32), ("total_ms", C.c_double), ("calls", C.c_int64)]


class DeviceArray:
    """A device buffer owned through the context (one HIP runtime per process)."""

    def __init__(self, ctx, nbytes):
        self.ctx = ctx
        self.nbytes = int(nbytes)
        p = C.c_void_p()
        _check(lib().phz_device_malloc(ctx.h, self.nbytes, C.byref(p)))
        self.ptr = p.value

    def upload(self, a):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        _check(lib().phz_memcpy(self.ctx.h, C.c_void_p(self.ptr), C.c_void_p(a.ctypes.data),
                                a.nbytes, 1))
        return self

    def download(self, shape, dtype):
        out = np.zeros(shape, dtype)
        assert out.nbytes <= self.nbytes
        _check(lib().phz_memcpy(self.ctx.h, C.c_void_p(out.ctypes.data), C.c_void_p(self.ptr),
                                out.nbytes, 2))
        return out

    def free(self):
        if self.ptr:
            lib().phz_device_free(self.ctx.h, C.c_void_p(self.ptr))
            self.ptr = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                               "(phaser_amd has no CPU fallback)")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in EXPORTS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


class PhaserError(RuntimeError):
    pass


def _check(rc):
    if rc != 0:
        raise PhaserError(f"phaser_amd error {rc}: {lib().phz_last_error().decode()}")


def _p(a):
    return None if a is None else C.c_void_p(a.ctypes.data)


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def default_config(**kw):
    cfg = Config()
    lib().phz_default_config(C.byref(cfg))
    for k, v in kw.items():
        if not hasattr(cfg, k):
            raise KeyError(k)
        setattr(cfg, k, v)
    return cfg


class Context:
    """Owns one phz_ctx (device, stream, resident tables, workspaces)."""

    def __init__(self, **kw):
        self.cfg = default_config(**kw)
        h = C.c_void_p()
        _check(lib().phz_create(C.byref(self.cfg), C.byref(h)))
        self.h = h
        self.B = self.cfg.bandwidth
        self.n = 2 * self.B  # S^2 grid
        self.no = 2 * (self.B + self.cfg.spherical_zero_padding)  # SO(3) grid
        self.N = self.cfg.n_voxels
        self.Np = self.N + 2 * self.cfg.spatial_zero_padding  # translation grid

    def close(self):
        if self.h:
            lib().phz_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # ---- end to end
    def register(self, target, source, grids=False, registered=False, target_extra=None,
                 source_extra=None):
        """target_extra / source_extra: (P, 2) reflectivity and ambient per
        point (a context with channels=4), or None (zero channels)."""
        t, s = _f32(target), _f32(source)
        te = None if target_extra is None else _f32(target_extra)
        se = None if source_extra is None else _f32(source_extra)
        for e, cl in ((te, t), (se, s)):
            assert e is None or e.shape == (cl.shape[0], 2)
        res = Result()
        rc = np.zeros(self.no ** 3, np.float32) if grids else None
        tc = np.zeros(self.Np ** 3, np.float32) if grids else None
        reg = np.zeros_like(s) if registered else None
        _check(lib().phz_register_channels(self.h, _p(t), _p(te), t.shape[0], _p(s), _p(se),
                                           s.shape[0], C.byref(res), _p(rc), _p(tc), _p(reg)))
        out = res.as_dict()
        if grids:
            out["rot_corr"] = rc
            out["trans_corr"] = tc
        if registered:
            out["registered"] = reg
        return out

    def register_device(self, d_target, Pt, d_source, Ps, d_rot=0, d_trans=0, d_reg=0, stream=0):
        res = Result()
        _check(lib().phz_register_device(self.h, C.c_void_p(d_target), Pt, C.c_void_p(d_source),
                                         Ps, C.byref(res), C.c_void_p(d_rot or None),
                                         C.c_void_p(d_trans or None), C.c_void_p(d_reg or None),
                                         C.c_void_p(stream or None)))
        return res.as_dict()

    def register_pairs_device(self, d_targets, Pt, d_sources, Ps, shared_target=False,
                              d_rot=0, d_trans=0, stream=0):
        n = len(d_sources)
        T = (C.c_void_p * len(d_targets))(*d_targets)
        S = (C.c_void_p * n)(*d_sources)
        PT = (C.c_int64 * len(Pt))(*Pt)
        PS = (C.c_int64 * n)(*Ps)
        res = (Result * n)()
        _check(lib().phz_register_pairs_device(self.h, n, int(shared_target), T, PT, S, PS, res,
                                               C.c_void_p(d_rot or None),
                                               C.c_void_p(d_trans or None),
                                               C.c_void_p(stream or None)))
        return results_as_dicts(res, n)

    def register_pairs_channels_device(self, d_targets, d_targets_ra, Pt, d_sources,
                                       d_sources_ra, Ps, shared_target=False, d_rot=0,
                                       d_trans=0, stream=0):
        """As register_pairs_device with one (reflectivity, ambient) device
        array per cloud (0 for none), on a context with channels=4."""
        n = len(d_sources)
        T = (C.c_void_p * len(d_targets))(*d_targets)
        TR = (C.c_void_p * len(d_targets))(*[p or None for p in d_targets_ra])
        S = (C.c_void_p * n)(*d_sources)
        SR = (C.c_void_p * n)(*[p or None for p in d_sources_ra])
        PT = (C.c_int64 * len(Pt))(*Pt)
        PS = (C.c_int64 * n)(*Ps)
        res = (Result * n)()
        _check(lib().phz_register_pairs_channels_device(
            self.h, n, int(shared_target), T, TR, PT, S, SR, PS, res, C.c_void_p(d_rot or None),
            C.c_void_p(d_trans or None), C.c_void_p(stream or None)))
        return results_as_dicts(res, n)

    def register_pairs(self, targets, sources, shared_target=False):
        """Host clouds; targets holds one cloud when shared_target is set."""
        tg = [np.ascontiguousarray(t, dtype=np.float32) for t in targets]
        sr = [np.ascontiguousarray(s, dtype=np.float32) for s in sources]
        n = len(sr)
        T = (C.c_void_p * len(tg))(*[t.ctypes.data for t in tg])
        S = (C.c_void_p * n)(*[s.ctypes.data for s in sr])
        PT = (C.c_int64 * len(tg))(*[len(t) for t in tg])
        PS = (C.c_int64 * n)(*[len(s) for s in sr])
        res = (Result * n)()
        _check(lib().phz_register_pairs(self.h, n, int(shared_target), T, PT, S, PS, res))
        return results_as_dicts(res, n)

    def memory_usage(self):
        b = C.c_size_t()
        _check(lib().phz_memory_usage(self.h, C.byref(b)))
        return b.value

    # ---- device memory and statistics
    def device_array(self, a):
        a = np.ascontiguousarray(a)
        return DeviceArray(self, a.nbytes).upload(a)

    def device_empty(self, nbytes):
        return DeviceArray(self, nbytes)

    def synchronize(self):
        _check(lib().phz_synchronize(self.h))

    def copy_kernel(self, d_dst, d_src, nbytes, repeats=1):
        """Mean ms per launch of the streaming copy kernel."""
        ms = C.c_float()
        _check(lib().phz_copy_kernel(self.h, C.c_void_p(d_dst), C.c_void_p(d_src), nbytes,
                                     repeats, C.byref(ms)))
        return ms.value

    def autotune(self, repeats=5):
        t = Tuning()
        _check(lib().phz_autotune(self.h, repeats, C.byref(t)))
        return t.as_dict()

    def tuning(self):
        t = Tuning()
        _check(lib().phz_get_tuning(self.h, C.byref(t)))
        return t.as_dict()

    def set_tuning(self, fwd_x=0, fwd_z=0, inv_x=0, plane_fwd=0, nn_lanes=0, fused_reg=0,
                   so3_reg=0, occ_fwd=0, inv_plane=0, nn_xcd=0, nn_seed=0, xz_slab=0):
        t = Tuning(fwd_x, fwd_z, inv_x, plane_fwd, nn_lanes, fused_reg, so3_reg, occ_fwd,
                   inv_plane, nn_xcd, nn_seed, xz_slab)
        _check(lib().phz_set_tuning(self.h, C.byref(t)))

    def set_graphs(self, on=True):
        _check(lib().phz_set_graphs(self.h, int(on)))

    def set_profiling(self, on=True):
        _check(lib().phz_set_profiling(self.h, int(on)))

    def stage_times(self):
        n = C.c_int()
        buf = (StageTime * 32)()
        _check(lib().phz_get_stage_times(self.h, buf, 32, C.byref(n)))
        return {buf[i].name.decode(): (buf[i].total_ms, buf[i].calls) for i in range(n.value)}

    # ---- stages
    def sample_s2(self, pts):
        pts = _f32(pts)
        out = np.zeros((2, self.n * self.n), np.float32)
        _check(lib().phz_sample_s2(self.h, _p(pts), pts.shape[0], _p(out)))
        return out

    def s2_transform(self, grids):
        g = _f32(grids).reshape(-1, self.n * self.n)
        out = np.zeros((g.shape[0], self.B * self.B), np.complex64)
        _check(lib().phz_s2_transform(self.h, g.shape[0], _p(g), _p(out)))
        return out

    def so3_correlate_coef(self, f, h, synthesis=False):
        f = np.ascontiguousarray(f, np.complex64)
        h = np.ascontiguousarray(h, np.complex64)
        corr = np.zeros(self.n ** 3, np.float32)
        syn = np.zeros(self.B * (2 * self.B - 1) * self.n, np.complex64) if synthesis else None
        am = C.c_uint32()
        _check(lib().phz_so3_correlate_coef(self.h, _p(f), _p(h), _p(corr), _p(syn), C.byref(am)))
        return (corr, int(am.value), syn) if synthesis else (corr, int(am.value))

    def spherical_correlate(self, f_grids, h_grids):
        f = _f32(f_grids)
        h = _f32(h_grids)
        Cn = f.shape[0]
        corr = np.zeros(self.n ** 3, np.float32)
        am = C.c_uint32()
        _check(lib().phz_spherical_correlate(self.h, Cn, _p(f), _p(h), _p(corr), C.byref(am)))
        return corr, int(am.value)

    def voxelize(self, pts):
        pts = _f32(pts)
        out = np.zeros((2, self.N ** 3), np.float32)
        cnt = np.zeros(self.N ** 3, np.float32)
        _check(lib().phz_voxelize(self.h, _p(pts), pts.shape[0], _p(out), _p(cnt)))
        return out, cnt

    def spatial_correlate(self, f, g):
        f = _f32(f)
        g = _f32(g)
        corr = np.zeros(self.N ** 3, np.float32)
        am = C.c_uint32()
        _check(lib().phz_spatial_correlate(self.h, f.shape[0], _p(f), _p(g), _p(corr), C.byref(am)))
        return corr, int(am.value)

    def spatial_correlate_single(self, f, g, N):
        f = _f32(f).reshape(-1)
        g = _f32(g).reshape(-1)
        corr = np.zeros(N ** 3, np.float32)
        am = C.c_uint32()
        _check(lib().phz_spatial_correlate_single(self.h, N, _p(f), _p(g), _p(corr), C.byref(am)))
        return corr, int(am.value)

    def spatial_correlate_low_pass(self, f, g, N, lower, upper, pad=0):
        f = _f32(f).reshape(-1)
        g = _f32(g).reshape(-1)
        corr = np.zeros((N + 2 * pad) ** 3, np.float32)
        am = C.c_uint32()
        _check(lib().phz_spatial_correlate_low_pass(self.h, N, pad, lower, upper, _p(f), _p(g),
                                                    _p(corr), C.byref(am)))
        return corr, int(am.value)

    def spherical_correlate_low_pass(self, f_grid, h_grid, lower, upper):
        f = _f32(f_grid).reshape(-1)
        h = _f32(h_grid).reshape(-1)
        corr = np.zeros(self.n ** 3, np.float32)
        am = C.c_uint32()
        _check(lib().phz_spherical_correlate_low_pass(self.h, _p(f), _p(h), lower, upper,
                                                      _p(corr), C.byref(am)))
        return corr, int(am.value)

    def mfma_peak(self, dtype=0, iters=2000):
        """Measured dense MFMA TFLOP/s (0 f16, 1 bf16, 2 f32)."""
        t = C.c_double()
        _check(lib().phz_mfma_peak(self.h, dtype, iters, C.byref(t)))
        return t.value

    def voxel_grid_filter(self, pts, leaf=0.25):
        """VoxelGridCmd: the voxel centroids, ordered by voxel index."""
        a = np.array(_f32(pts).reshape(-1, 4))
        n = C.c_int64()
        _check(lib().phz_voxel_grid_filter(self.h, _p(a), a.shape[0], leaf, C.byref(n)))
        return a[: n.value].copy()

    def pass_through_filter(self, pts, field=2, lower=-5.0, upper=3.0, negative=True):
        """PassThroughGndFilterCmd's defaults: drop z in [-5, 3]."""
        a = np.array(_f32(pts).reshape(-1, 4))
        n = C.c_int64()
        _check(lib().phz_pass_through_filter(self.h, _p(a), a.shape[0], field, lower, upper,
                                             int(negative), C.byref(n)))
        return a[: n.value].copy()

    def zscore_peaks(self, corr, max_peaks=1 << 20, **kw):
        """ZScorePeakExtraction on the device; kw: lag_percentile, lag_max,
        threshold, influence, filter_threshold (the flags' defaults
        otherwise).  Returns the signal indices."""
        p = ZScoreParams()
        lib().phz_zscore_default_params(C.byref(p))
        for k, v in kw.items():
            setattr(p, k, v)
        c = _f32(corr).reshape(-1)
        out = np.zeros(max(max_peaks, 1), np.uint32)
        cnt = C.c_int64()
        _check(lib().phz_zscore_peaks(self.h, _p(c), c.size, C.byref(p), _p(out), max_peaks,
                                      C.byref(cnt)))
        return out[: min(cnt.value, max_peaks)].copy()

    def spatial_correlate_single_padded(self, f, g, N, pad):
        f = _f32(f).reshape(-1)
        g = _f32(g).reshape(-1)
        corr = np.zeros((N + 2 * pad) ** 3, np.float32)
        am = C.c_uint32()
        _check(lib().phz_spatial_correlate_single_padded(self.h, N, pad, _p(f), _p(g), _p(corr),
                                                         C.byref(am)))
        return corr, int(am.value)

    def spatial_correlate_single_noalloc(self, d_f, d_g, d_corr, N):
        """phz_spatial_correlate_single_noalloc on device buffers (DeviceArray
        or raw addresses; d_corr may be None); returns the argmax."""
        addr = lambda a: None if a is None else C.c_void_p(getattr(a, "ptr", a))
        am = C.c_uint32()
        _check(lib().phz_spatial_correlate_single_noalloc(self.h, N, addr(d_f), addr(d_g),
                                                          addr(d_corr), C.byref(am), None))
        return int(am.value)
