"""CPU-side checks of the drop-in boundary: the HIP library loads and exports
every entry point declared in include/*.h (no compute calls: no GPU here),
and the ctypes structures match the C layout."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared(header, prefix="phz_"):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(" + prefix + r"[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    from phaser_amd import capi
    if not os.path.exists(capi.LIB_PATH):
        pytest.fail("libphaser_amd.so is not built")
    L = C.CDLL(capi.LIB_PATH)
    names = _declared("phaser_amd.h")
    assert len(names) >= 15
    for n in names:
        assert hasattr(L, n), n
        assert n in capi.EXPORTS, f"{n} has no ctypes binding"


def test_ply_reader_exports_every_declared_symbol():
    from phaser_amd import ply
    L = ply.lib()
    names = _declared("phaser_ply.h")
    assert len(names) == 5
    for n in names:
        assert hasattr(L, n), n
        assert n in ply.EXPORTS, f"{n} has no ctypes binding"


def test_synth_library_exports():
    from phaser_amd import synth
    L = synth._load()
    for n in _declared("phaser_synth.h"):
        assert hasattr(L, n), n


def test_struct_layouts():
    from phaser_amd import capi
    assert C.sizeof(capi.Config) == 4 * 8 + 8 + 4 * 12
    # rot/trans index + values (16) + zyz,quat,euler (80) + T (64) + translation (24)
    # + trans_cov (72) + rot_omega (32) + rot_axes (128) + status, reserved (8)
    assert C.sizeof(capi.Result) == 16 + 80 + 64 + 24 + 72 + 32 + 128 + 8


def test_default_config_and_error_path_without_gpu():
    from phaser_amd import capi
    cfg = capi.default_config()
    assert cfg.bandwidth == 64 and cfg.n_voxels == 128
    assert cfg.discretize_lower == -50 and cfg.discretize_upper == 50
    bad = capi.default_config(bandwidth=161)  # the search grid holds 2 * 160 bands
    h = C.c_void_p()
    rc = capi.lib().phz_create(C.byref(bad), C.byref(h))
    assert rc == -1
    assert b"bandwidth" in capi.lib().phz_last_error()
    bad2 = capi.default_config(bandwidth=150, spherical_zero_padding=11)
    assert capi.lib().phz_create(C.byref(bad2), C.byref(h)) == -1
    bad3 = capi.default_config(n_voxels=257)
    assert capi.lib().phz_create(C.byref(bad3), C.byref(h)) == -1
    # any bandwidth and voxel count in range is accepted by the checks (the
    # estimate runs them without a device)
    for kw in (dict(bandwidth=150, spherical_zero_padding=10, n_voxels=201),
               dict(bandwidth=7, n_voxels=67)):
        b = C.c_size_t()
        assert capi.lib().phz_estimate_memory(C.byref(capi.default_config(**kw)), C.byref(b)) == 0
        assert b.value > 0


def test_serve_library_exports_every_declared_symbol():
    from phaser_amd import serve
    if not os.path.exists(serve.LIB_PATH):
        pytest.fail("libphaser_serve.so is not built")
    L = serve.lib()
    names = _declared("phaser_serve.h", "phs_")
    assert len(names) >= 9
    for n in names:
        assert hasattr(L, n), n
        assert n in serve.EXPORTS, f"{n} has no ctypes binding"


_LAYOUT_C = r"""
#include <stdio.h>
#include <stddef.h>
#include "phaser_amd.h"
#include "phaser_serve.h"
#include "phaser_ply.h"
#define S(t) printf(#t " %zu\n", sizeof(t))
#define O(t, f) printf(#t "." #f " %zu\n", offsetof(t, f))
int main(void) {
  S(phz_config); S(phz_result); S(phz_stage_time); S(phz_tuning);
  S(phz_ply_fields); S(phz_ply_cloud); O(phz_ply_cloud, range); O(phz_ply_cloud, has_range);
  O(phz_tuning, plane_fwd); O(phz_tuning, nn_lanes); O(phz_tuning, fused_reg);
  O(phz_tuning, so3_reg); O(phz_tuning, occ_fwd); O(phz_tuning, inv_plane); O(phz_tuning, nn_xcd);
  O(phz_tuning, nn_seed);
  O(phz_tuning, ms_fwd_x); O(phz_tuning, ms_plane);
  O(phz_tuning, ms_fused); O(phz_tuning, ms_so3);
  S(phs_config); S(phs_request); S(phs_response); S(phs_stats);
  O(phz_config, max_points); O(phz_config, device); O(phz_result, zyz);
  O(phz_result, T_rot); O(phz_result, translation); O(phz_result, trans_cov);
  O(phz_result, rot_omega); O(phz_result, rot_axes); O(phz_result, status);
  O(phz_config, bingham_peak_neighbors); O(phz_config, gaussian_peak_neighbors);
  O(phz_config, channels); O(phz_config, method); O(phz_config, skip_rot_axes);
  O(phs_config, bingham_peak_neighbors); O(phs_config, gaussian_peak_neighbors);
  O(phs_config, method); O(phs_config, skip_rot_axes);
  O(phs_config, max_points); O(phs_config, hbm_budget); O(phs_config, batch_window_ms);
  O(phs_config, precision); O(phs_config, graphs); O(phz_config, precision);
  O(phs_request, target); O(phs_request, Ps); O(phs_request, tag);
  O(phs_response, result); O(phs_response, latency_ms); O(phs_response, error);
  O(phs_stats, contexts_live); O(phs_stats, queued);
  return 0;
}
"""


def test_ctypes_layouts_match_the_c_headers(tmp_path):
    """Sizes and field offsets of every ctypes structure against the C
    compiler's view of include/*.h."""
    from phaser_amd import capi, ply, serve
    src = tmp_path / "layout.c"
    src.write_text(_LAYOUT_C)
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), str(src), "-o",
                    str(exe)], check=True)
    got = dict(line.split() for line in subprocess.run([str(exe)], capture_output=True,
                                                       text=True, check=True).stdout.splitlines())
    types = {"phz_config": capi.Config, "phz_result": capi.Result,
             "phz_stage_time": capi.StageTime, "phz_tuning": capi.Tuning,
             "phs_config": serve.ServeConfig,
             "phs_request": serve.Request, "phs_response": serve.Response,
             "phs_stats": serve.Stats, "phz_ply_fields": ply.Fields, "phz_ply_cloud": ply.Cloud}
    for key, val in got.items():
        if "." in key:
            t, f = key.split(".")
            assert getattr(types[t], f).offset == int(val), key
        else:
            assert C.sizeof(types[key]) == int(val), key


def test_memory_estimate_of_the_optional_buffers():
    """The extra device bytes of four channels (per-point pairs, two voxel
    sums, a second spectra set, twice the S^2 grids) and of the peak windows
    (one rotation and one translation grid per pair), from the estimate."""
    from phaser_amd import capi

    def est(**kw):
        cfg = capi.default_config(bandwidth=32, n_voxels=64, max_points=200000, max_batch=4,
                                  **kw)
        b = C.c_size_t()
        assert capi.lib().phz_estimate_memory(C.byref(cfg), C.byref(b)) == 0
        return b.value
    B, n, N, P, M, cap = 32, 64, 64, 4, 8, 200000
    nq, T3, H = n * n, N ** 3, N * N * (N // 2 + 1)
    base = est()
    grids = (max(4 * M, 16) - max(2 * M, 16)) * (nq * 4 + B * n * 8 + B * B * 8)
    assert est(channels=4) - base == M * cap * 8 + 2 * M * T3 * 4 + max(2 * M, 16) * H * 8 + grids
    assert est(bingham_peak_neighbors=2, gaussian_peak_neighbors=1) - base == \
        P * n ** 3 * 4 + P * T3 * 4


def test_memory_estimate_without_gpu():
    """phz_estimate_memory validates like phz_create and makes no device call."""
    from phaser_amd import capi
    est = {}
    for B, N, P in ((32, 64, 1), (64, 128, 1), (64, 128, 16), (128, 256, 1)):
        cfg = capi.default_config(bandwidth=B, n_voxels=N, max_batch=P)
        b = C.c_size_t()
        assert capi.lib().phz_estimate_memory(C.byref(cfg), C.byref(b)) == 0
        est[(B, N, P)] = b.value
    # at least the voxel accumulators and spectra of the two clouds
    assert est[(64, 128, 1)] > 16 * 128 ** 3 * 4
    assert est[(32, 64, 1)] < est[(64, 128, 1)] < est[(64, 128, 16)]
    assert est[(64, 128, 1)] < est[(128, 256, 1)]
    bad = capi.default_config(bandwidth=161)
    b = C.c_size_t()
    assert capi.lib().phz_estimate_memory(C.byref(bad), C.byref(b)) == -1


def test_results_as_dicts_equals_as_dict():
    """The column-wise conversion of a batch of results gives what
    Result.as_dict() gives for each: the same keys, types, shapes and values."""
    from phaser_amd import capi
    rng = np.random.default_rng(0)
    n = 5
    res = (capi.Result * n)()
    for k in range(n):
        r = res[k]
        r.rot_index, r.trans_index, r.status = 7 * k + 1, 11 * k + 2, k % 3
        r.rot_value, r.trans_value = float(rng.normal()), float(rng.normal() * 1e9)
        for name, ct in capi.Result._fields_:
            if hasattr(ct, "_length_"):
                arr = getattr(r, name)
                for i in range(ct._length_):
                    arr[i] = float(rng.normal())
    got = capi.results_as_dicts(res, n)
    assert len(got) == n
    for a, b in zip(got, [r.as_dict() for r in res]):
        assert list(a) == list(b)
        for key, vb in b.items():
            va = a[key]
            if isinstance(vb, np.ndarray):
                assert va.dtype == vb.dtype and va.shape == vb.shape, key
                np.testing.assert_array_equal(va, vb)
            else:
                assert type(va) is type(vb) and va == vb, key
    two = capi.results_as_dicts(res, 2)
    assert [d["rot_index"] for d in two] == [d["rot_index"] for d in got[:2]]
