"""The request loop on the GPU: every result of a pair served in a mix of
shapes, batches and shared targets equals the result of the same pair
registered alone (phz_register on a context of its shape).

Tolerance (alone vs in the mix): rotation bin equal and |C|^2 equal to 1e-6
relative (the SO(3) path is deterministic); translation voxel equal unless
the two correlation peaks differ by less than 1e-4 relative (voxel sums are
FP32 atomics whose order varies), translation value within 1e-4 relative,
rotation matrix within 1e-6 and translation within 1e-6 m of the alone pose
whenever the bins agree."""
import random
import time

import numpy as np
import pytest

from phaser_amd import capi, serve, synth

pytestmark = pytest.mark.gpu

SHAPES = [(32, 64), (48, 96), (64, 128)]


def _close(a, b, rtol):
    return abs(a - b) <= rtol * max(abs(a), abs(b), 1e-30)


def _same(alone, got):
    assert got["status"] == 0, got["error"]
    assert got["rot_index"] == alone["rot_index"]
    assert _close(got["rot_value"], alone["rot_value"], 1e-6)
    if got["trans_index"] != alone["trans_index"]:
        assert _close(got["trans_value"], alone["trans_value"], 1e-4), "translation bin differs"
        return
    assert _close(got["trans_value"], alone["trans_value"], 1e-4)
    np.testing.assert_allclose(got["T_rot"], alone["T_rot"], atol=1e-6)
    np.testing.assert_allclose(got["translation"], alone["translation"], atol=1e-6)


def test_served_mix_equals_alone():
    rng = random.Random(5)
    work = []  # (shape, target, source, key)
    for (B, N) in SHAPES:
        for seed in range(4):
            src, tgt, _, _ = synth.mixed_pair(seed, B)
            work.append(((B, N), tgt, src, 0))
    # one target against several sources, sharing its spectra
    tgt0 = synth.mixed_pair(100, 64)[1]
    for seed in range(101, 107):
        work.append(((64, 128), tgt0, synth.mixed_pair(seed, 64)[0], 42))
    rng.shuffle(work)
    print(f"serve mix: {len(work)} pairs over {len(SHAPES)} shapes", flush=True)

    alone = []
    ctxs = {}
    try:
        for (shape, tgt, src, _) in work:
            if shape not in ctxs:
                ctxs[shape] = capi.Context(bandwidth=shape[0], n_voxels=shape[1])
            alone.append(ctxs[shape].register(tgt, src))
    finally:
        for c in ctxs.values():
            c.close()

    with serve.Server(max_batch=8, contexts_per_shape=2, hbm_budget=48 << 30) as srv:
        tickets = []
        for i, (shape, tgt, src, key) in enumerate(work):
            tickets.append(srv.submit(tgt, src, bandwidth=shape[0], n_voxels=shape[1],
                                      target_key=key, tag=i))
            if i % 5 == 4:
                time.sleep(rng.uniform(0.0, 0.004))  # arrivals while batches run
        got = [srv.wait(t) for t in tickets]
        st = srv.stats()
    print(f"serve mix stats: {st}", flush=True)
    for i, (a, g) in enumerate(zip(alone, got)):
        assert g["tag"] == i
        _same(a, g)
    assert st["completed"] == len(work) and st["failed"] == 0
    assert st["batches"] < len(work), "no request shared a batch"
    assert any(g["shared_target"] for g in got)
    assert st["bytes_peak"] <= 48 << 30


def test_served_under_tight_budget():
    """A budget that holds one context: shapes take turns, results unchanged."""
    work = []
    for (B, N) in SHAPES[:2]:
        for seed in range(3):
            src, tgt, _, _ = synth.mixed_pair(seed + 10, B)
            work.append(((B, N), tgt, src))
    need = []
    for (B, N) in SHAPES[:2]:
        cfg = capi.default_config(bandwidth=B, n_voxels=N, max_batch=4)
        import ctypes as C
        b = C.c_size_t()
        assert capi.lib().phz_estimate_memory(C.byref(cfg), C.byref(b)) == 0
        need.append(b.value)
    budget = max(need)
    alone = []
    for (shape, tgt, src) in work:
        with capi.Context(bandwidth=shape[0], n_voxels=shape[1]) as c:
            alone.append(c.register(tgt, src))
    with serve.Server(max_batch=4, contexts_per_shape=2, hbm_budget=budget) as srv:
        tickets = [srv.submit(t, s, bandwidth=sh[0], n_voxels=sh[1]) for (sh, t, s) in work]
        got = [srv.wait(t) for t in tickets]
        st = srv.stats()
    for a, g in zip(alone, got):
        _same(a, g)
    assert st["bytes_peak"] <= budget


def test_failed_peak_window_fails_only_its_request():
    """A pair whose Bingham window would end at (2B)^3 (the reference's
    CHECK_LT(end, n), bingham-peak-based-eval.cc:69) fails with -4 and
    status PHZ_PAIR_ROT_WINDOW; the other pairs served with it succeed
    and equal their registration alone."""
    B, N = 32, 64
    pairs = [synth.mixed_pair(seed, B) for seed in (20, 21, 22)]
    with capi.Context(bandwidth=B, n_voxels=N) as c:
        idx = [c.register(t, s)["rot_index"] for (s, t, _, _) in pairs]
    assert idx[0] not in idx[1:], "pick seeds with distinct rotation peaks"
    k = (2 * B) ** 3 - idx[0]
    print(f"serve peak window: argmax {idx}, bingham_peak_neighbors {k}", flush=True)
    with capi.Context(bandwidth=B, n_voxels=N, bingham_peak_neighbors=k) as c:
        alone = [c.register(t, s) for (s, t, _, _) in pairs[1:]]
    with serve.Server(max_batch=4, batch_window_ms=5.0, peak_neighbors=(k, 0)) as srv:
        tickets = [srv.submit(t, s, bandwidth=B, n_voxels=N, tag=i)
                   for i, (s, t, _, _) in enumerate(pairs)]
        got = [srv.wait(t) for t in tickets]
        st = srv.stats()
    assert got[0]["status"] == -4 and "status 1" in got[0]["error"], got[0]
    for a, g in zip(alone, got[1:]):
        _same(a, g)
        np.testing.assert_allclose(g["rot_omega"], a["rot_omega"], atol=1e-12)
    assert st["failed"] == 1 and st["completed"] == 2


def test_served_sph_method_equals_alone():
    """A server whose contexts run SphRegistration (method 1, refine_rot_z
    false) serves each pair as a context of that method registers it alone."""
    B, N = 32, 64
    pairs = [synth.mixed_pair(seed, B) for seed in (30, 31, 32, 33)]
    with capi.Context(bandwidth=B, n_voxels=N, method=1, skip_rot_axes=4) as c:
        alone = [c.register(t, s) for (s, t, _, _) in pairs]
    print("served sph method", flush=True)
    with serve.Server(max_batch=4, batch_window_ms=5.0, method=1, skip_rot_axes=4) as srv:
        tickets = [srv.submit(t, s, bandwidth=B, n_voxels=N) for (s, t, _, _) in pairs]
        got = [srv.wait(t) for t in tickets]
    for a, g in zip(alone, got):
        _same(a, g)
        np.testing.assert_array_equal(g["quat"], a["quat"])


@pytest.mark.parametrize("precision", [0, 1])
def test_served_full_mix_both_forms_equals_alone(precision):
    """The serving configuration's whole bandwidth mix, B in {32, 64, 96, 128}
    with 64^3 .. 256^3 voxels, in the plain FP32 form and the split-FP16
    MFMA form: every served pair equals its registration alone in the same
    form."""
    shapes = [(32, 64), (64, 128), (96, 192), (128, 256)]
    rng = random.Random(11 + precision)
    work = []
    for (B, N) in shapes:
        for seed in range(3):
            src, tgt, _, _ = synth.mixed_pair(200 + seed, B)
            work.append(((B, N), tgt, src))
    rng.shuffle(work)
    print(f"serve full mix precision={precision}: {len(work)} pairs", flush=True)
    alone = []
    ctxs = {}
    try:
        for (shape, tgt, src) in work:
            if shape not in ctxs:
                ctxs[shape] = capi.Context(bandwidth=shape[0], n_voxels=shape[1],
                                           precision=precision)
            alone.append(ctxs[shape].register(tgt, src))
    finally:
        for c in ctxs.values():
            c.close()
    with serve.Server(max_batch=4, contexts_per_shape=1, hbm_budget=96 << 30,
                      precision=precision) as srv:
        tickets = []
        for i, (shape, tgt, src) in enumerate(work):
            tickets.append(srv.submit(tgt, src, bandwidth=shape[0], n_voxels=shape[1], tag=i))
            if i % 3 == 2:
                time.sleep(rng.uniform(0.0, 0.003))
        got = [srv.wait(t) for t in tickets]
        st = srv.stats()
    print(f"serve full mix stats: {st}", flush=True)
    for a, g in zip(alone, got):
        _same(a, g)
    assert st["completed"] == len(work) and st["failed"] == 0
