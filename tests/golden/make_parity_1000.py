"""Writes the oracle's argmax fixtures for the argmax-match rate at B=64 with
128^3 voxels over seeds 0..999 (the FP64 oracle takes ~6 s per pair, too
slow to recompute inside a GPU test run).

Workload: the seeded generator at 128 rings x 1024 azimuths (~80k points),
target cropped to 70% overlap, noise sigma 2 cm, all flags at their defaults
(SphOptRegistration, no padding, no peak neighbours).  target_mode 0: the
target is the source's points rigidly moved; target_mode 1: the target is
re-cast from the moved sensor.

Per seed: both argmax indices and values, the indices and values of the
eight largest entries of each grid (so that a device argmax that differs
can be checked to be a near-tie of the oracle's own grid), the grids' max
|value|, the pose, and the point counts.

Usage: python tests/golden/make_parity_1000.py [mode ...] [--threads T]
"""
import argparse
import concurrent.futures as cf
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

SEEDS = 1000
TOP = 8


def fixture_path(mode):
    return os.path.join(HERE, f"parity_b64_n128_mode{mode}.npz")


def one(seed, mode):
    from oracle import oracle as po
    from phaser_amd import synth
    src, tgt, _, _ = synth.pair(seed=seed, target_mode=mode, **synth.CONFIG_B64)
    r = po.register(tgt, src, bw=64, n_voxels=128, want_grids=True)
    rc, tc = r["rot_corr"], r["trans_corr"]
    ri = np.argpartition(rc, -TOP)[-TOP:]
    ri = ri[np.argsort(-rc[ri], kind="stable")]
    ti = np.argpartition(tc, -TOP)[-TOP:]
    ti = ti[np.argsort(-tc[ti], kind="stable")]
    return dict(rot_index=r["rot_index"], trans_index=r["trans_index"],
                rot_value=r["rot_value"], trans_value=r["trans_value"],
                rot_top_index=ri, rot_top_value=rc[ri], trans_top_index=ti,
                trans_top_value=tc[ti], rot_max_abs=np.max(np.abs(rc)),
                trans_max_abs=np.max(np.abs(tc)), quat=r["quat"], translation=r["translation"],
                n_points=np.array([len(src), len(tgt)]))


def make(mode, threads):
    recs = [None] * SEEDS
    t0 = time.time()
    with cf.ThreadPoolExecutor(max_workers=threads) as ex:
        futs = {ex.submit(one, s, mode): s for s in range(SEEDS)}
        for k, f in enumerate(cf.as_completed(futs)):
            recs[futs[f]] = f.result()
            if k % 50 == 49:
                print(f"mode {mode}: {k + 1}/{SEEDS} pairs, {time.time() - t0:.0f} s", flush=True)
    out = {"seeds": np.arange(SEEDS), "target_mode": np.array(mode)}
    for key in recs[0]:
        out[key] = np.stack([np.asarray(r[key]) for r in recs])
    np.savez_compressed(fixture_path(mode), **out)
    print(f"wrote {fixture_path(mode)}", flush=True)


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("modes", nargs="*", type=int, default=[1, 0])
    ap.add_argument("--threads", type=int, default=os.cpu_count() or 1)
    a = ap.parse_args()
    for m in a.modes:
        make(m, a.threads)
