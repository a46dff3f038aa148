"""Seeded synthetic scan pairs (ctypes wrapper over libphaser_synth.so).

Host-side workload generator shared by tests, the benchmark and the CPU
baseline; it is not part of the GPU path.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "lib", "libphaser_synth.so")
_lib = None


class SynthConfig(C.Structure):
    _fields_ = [("seed", C.c_uint64), ("rings", C.c_int32), ("azimuths", C.c_int32),
                ("elev_min_deg", C.c_float), ("elev_max_deg", C.c_float),
                ("min_range", C.c_float), ("max_range", C.c_float),
                ("ground_depth", C.c_float), ("n_boxes", C.c_int32),
                ("n_cylinders", C.c_int32), ("scene_radius", C.c_float),
                ("max_translation", C.c_float), ("noise_sigma", C.c_float),
                ("overlap", C.c_float), ("outlier_frac", C.c_float),
                ("random_rotation", C.c_int32), ("target_mode", C.c_int32)]


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError(f"{_LIB_PATH} is missing: run __graft_entry__.build()")
        L = C.CDLL(_LIB_PATH)
        L.phz_synth_default_config.argtypes = [C.POINTER(SynthConfig)]
        L.phz_synth_capacity.argtypes = [C.POINTER(SynthConfig)]
        L.phz_synth_capacity.restype = C.c_int64
        L.phz_synth_pair.argtypes = [C.POINTER(SynthConfig), C.c_void_p, C.c_int64,
                                     C.POINTER(C.c_int64), C.c_void_p, C.c_int64,
                                     C.POINTER(C.c_int64), C.c_void_p, C.c_void_p]
        L.phz_synth_pair.restype = C.c_int
        _lib = L
    return _lib


def config(**kw):
    cfg = SynthConfig()
    _load().phz_synth_default_config(C.byref(cfg))
    for k, v in kw.items():
        if not hasattr(cfg, k):
            raise KeyError(k)
        setattr(cfg, k, v)
    return cfg


def pair(seed=0, **kw):
    """Return (source, target, R, t): clouds are (P,4) float32 x,y,z,intensity
    and the target satisfies x_t ~= R x_s + t."""
    cfg = config(seed=seed, **kw)
    L = _load()
    cap = L.phz_synth_capacity(C.byref(cfg))
    src = np.zeros((cap, 4), np.float32)
    tgt = np.zeros((cap, 4), np.float32)
    ns, nt = C.c_int64(), C.c_int64()
    R = np.zeros(9)
    t = np.zeros(3)
    rc = L.phz_synth_pair(C.byref(cfg), src.ctypes.data, cap, C.byref(ns),
                          tgt.ctypes.data, cap, C.byref(nt), R.ctypes.data, t.ctypes.data)
    if rc != 0:
        raise RuntimeError(f"phz_synth_pair failed ({rc})")
    return src[: ns.value].copy(), tgt[: nt.value].copy(), R.reshape(3, 3), t


# workloads named by their sizes
CONFIG_B32 = dict(rings=64, azimuths=1024, noise_sigma=0.02)            # ~30k-40k points
CONFIG_B64 = dict(rings=128, azimuths=1024, noise_sigma=0.02, overlap=0.7)  # ~80k points
CONFIG_B128 = dict(rings=256, azimuths=1024, noise_sigma=0.02)          # ~200k points

# rings per bandwidth for the mixed batching and serving workloads
RINGS_FOR_BANDWIDTH = {32: 64, 48: 96, 64: 128, 96: 192, 128: 256}


def mixed_config(seed, bandwidth):
    """Batching/serving pair parameters: overlap U[40,90]%, noise sigma
    U[1,5] cm and 5% uniform outliers, drawn from the seed."""
    rng = np.random.default_rng([seed, bandwidth])
    return dict(rings=RINGS_FOR_BANDWIDTH[bandwidth], azimuths=1024,
                overlap=float(rng.uniform(0.4, 0.9)), noise_sigma=float(rng.uniform(0.01, 0.05)),
                outlier_frac=0.05)


def mixed_pair(seed, bandwidth):
    return pair(seed=seed, **mixed_config(seed, bandwidth))
