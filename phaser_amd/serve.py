"""ctypes binding of the request loop in include/phaser_serve.h.

A Server registers scan pairs of mixed bandwidth and voxel count as they are
submitted: per (bandwidth, n_voxels) it keeps phaser_amd contexts with
resident tables, batches whatever is queued when a context is free, and
admits contexts against a device-memory budget.  The point arrays of a
request are kept alive here until its response has been returned.
"""
import ctypes as C
import os

import numpy as np

from . import capi

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libphaser_serve.so")
_lib = None


class ServeConfig(C.Structure):
    _fields_ = [("device", C.c_int32), ("max_batch", C.c_int32),
                ("contexts_per_shape", C.c_int32), ("discretize_lower", C.c_int32),
                ("discretize_upper", C.c_int32), ("autotune", C.c_int32),
                ("max_points", C.c_int64), ("hbm_budget", C.c_uint64),
                ("batch_window_ms", C.c_double), ("precision", C.c_int32),
                ("graphs", C.c_int32), ("bingham_peak_neighbors", C.c_int32),
                ("gaussian_peak_neighbors", C.c_int32), ("method", C.c_int32),
                ("skip_rot_axes", C.c_int32)]


class Request(C.Structure):
    _fields_ = [("bandwidth", C.c_int32), ("n_voxels", C.c_int32), ("target_key", C.c_uint64),
                ("target", C.c_void_p), ("Pt", C.c_int64), ("source", C.c_void_p),
                ("Ps", C.c_int64), ("tag", C.c_uint64)]


class Response(C.Structure):
    _fields_ = [("ticket", C.c_uint64), ("tag", C.c_uint64), ("status", C.c_int32),
                ("batch_pairs", C.c_int32), ("shared_target", C.c_int32),
                ("reserved", C.c_int32), ("result", capi.Result), ("queue_ms", C.c_double),
                ("service_ms", C.c_double), ("latency_ms", C.c_double),
                ("error", C.c_char * 160)]

    def as_dict(self):
        d = self.result.as_dict() if self.status == 0 else {}
        d.update(ticket=int(self.ticket), tag=int(self.tag), status=int(self.status),
                 batch_pairs=int(self.batch_pairs), shared_target=bool(self.shared_target),
                 queue_ms=float(self.queue_ms), service_ms=float(self.service_ms),
                 latency_ms=float(self.latency_ms), error=self.error.decode())
        return d


class Stats(C.Structure):
    _fields_ = [(k, C.c_uint64) for k in ("submitted", "completed", "failed", "batches",
                                          "pairs_in_batches", "contexts_created",
                                          "contexts_released", "bytes_in_use", "bytes_peak")] + \
               [("contexts_live", C.c_int32), ("queued", C.c_int32)]


EXPORTS = {
    "phs_default_config": (None, [C.POINTER(ServeConfig)]),
    "phs_create": (C.c_int, [C.POINTER(ServeConfig), C.POINTER(C.c_void_p)]),
    "phs_destroy": (None, [C.c_void_p]),
    "phs_submit": (C.c_int, [C.c_void_p, C.POINTER(Request), C.POINTER(C.c_uint64)]),
    "phs_wait": (C.c_int, [C.c_void_p, C.c_uint64, C.POINTER(Response)]),
    "phs_poll": (C.c_int, [C.c_void_p, C.POINTER(Response), C.c_int, C.POINTER(C.c_int),
                           C.c_int]),
    "phs_drain": (C.c_int, [C.c_void_p]),
    "phs_get_stats": (C.c_int, [C.c_void_p, C.POINTER(Stats)]),
    "phs_last_error": (C.c_char_p, []),
}


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with __graft_entry__.build()")
        capi.lib()  # the kernels' library first (same directory, rpath $ORIGIN)
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in EXPORTS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


class ServeError(RuntimeError):
    def __init__(self, rc, msg):
        super().__init__(f"phaser_serve error {rc}: {msg}")
        self.rc = rc


def _check(rc):
    if rc != 0:
        raise ServeError(rc, lib().phs_last_error().decode())


class Server:
    def __init__(self, max_batch=16, contexts_per_shape=2, hbm_budget=64 << 30,
                 max_points=262144, batch_window_ms=0.0, device=0, bounds=(-50, 50),
                 autotune=False, precision=0, graphs=True, peak_neighbors=(0, 0), method=0,
                 skip_rot_axes=0):
        cfg = ServeConfig()
        lib().phs_default_config(C.byref(cfg))
        cfg.graphs = int(graphs)
        cfg.bingham_peak_neighbors, cfg.gaussian_peak_neighbors = peak_neighbors
        cfg.method, cfg.skip_rot_axes = method, skip_rot_axes
        cfg.max_batch = max_batch
        cfg.contexts_per_shape = contexts_per_shape
        cfg.hbm_budget = int(hbm_budget)
        cfg.max_points = max_points
        cfg.batch_window_ms = batch_window_ms
        cfg.autotune = int(autotune)
        cfg.precision = int(precision)
        cfg.device = device
        cfg.discretize_lower, cfg.discretize_upper = bounds
        h = C.c_void_p()
        _check(lib().phs_create(C.byref(cfg), C.byref(h)))
        self.h = h.value
        self._keep = {}

    def submit(self, target, source, bandwidth=64, n_voxels=128, target_key=0, tag=0):
        """Queues one registration (target = cloud_prev, source = cloud_cur)."""
        t = np.ascontiguousarray(target, dtype=np.float32)
        s = np.ascontiguousarray(source, dtype=np.float32)
        assert t.ndim == 2 and t.shape[1] == 4 and s.ndim == 2 and s.shape[1] == 4
        q = Request(bandwidth, n_voxels, target_key, t.ctypes.data, len(t), s.ctypes.data,
                    len(s), tag)
        ticket = C.c_uint64()
        _check(lib().phs_submit(self.h, C.byref(q), C.byref(ticket)))
        self._keep[ticket.value] = (t, s)
        return ticket.value

    def wait(self, ticket):
        r = Response()
        _check(lib().phs_wait(self.h, ticket, C.byref(r)))
        self._keep.pop(ticket, None)
        return r.as_dict()

    def poll(self, max_count=64, wait_ms=0):
        buf = (Response * max_count)()
        n = C.c_int()
        _check(lib().phs_poll(self.h, buf, max_count, C.byref(n), wait_ms))
        out = [buf[i].as_dict() for i in range(n.value)]
        for r in out:
            self._keep.pop(r["ticket"], None)
        return out

    def drain(self):
        _check(lib().phs_drain(self.h))

    def stats(self):
        st = Stats()
        _check(lib().phs_get_stats(self.h, C.byref(st)))
        return {k: getattr(st, k) for k, _ in Stats._fields_}

    def close(self):
        if self.h:
            lib().phs_destroy(self.h)
            self.h = None
            self._keep.clear()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
