"""Argmax-match sweep: registers the seeded pairs of an argmax fixture on the
GPU and compares both argmax indices with the FP64 oracle's recorded ones.

A fixture (written by tests/golden/make_parity_1000.py) holds, per seed, the
oracle's rotation bin and translation voxel, and the indices and values of
the eight largest entries of each of its grids.  A device index that differs
from the oracle's is classified as a near-tie when it is one of those eight
and the oracle's own values at the two indices differ by at most NEAR_TIE of
the grid's largest magnitude: FP32 cannot order such a pair (the oracle's
FP64 grid itself would reorder them under a last-bit change of the inputs).

Used by the GPU parity test and by the benchmark's argmax-match figure.
"""
import concurrent.futures as cf
import os

import numpy as np

from . import synth

NEAR_TIE = 1e-4


def load(path):
    with np.load(path, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def _classify(gpu_idx, fx, which, i):
    """None if gpu_idx equals the oracle's index, else the relative gap (or
    inf when the index is not among the oracle's top entries)."""
    o_idx = int(fx[f"{which}_index"][i])
    if gpu_idx == o_idx:
        return None
    top_i = fx[f"{which}_top_index"][i]
    top_v = fx[f"{which}_top_value"][i]
    hit = np.nonzero(top_i == gpu_idx)[0]
    if hit.size == 0:
        return float("inf")
    scale = float(fx[f"{which}_max_abs"][i])
    return float(top_v[0] - top_v[hit[0]]) / scale


def sweep(fx, bandwidth=64, n_voxels=128, batch=64, threads=16, precision=0, limit=None,
          context_kw=None):
    """Registers every fixture seed (or the first `limit`) in batches of
    `batch` pairs and returns the match statistics and the mismatches."""
    from .capi import Context
    seeds = fx["seeds"][:limit] if limit else fx["seeds"]
    mode = int(fx["target_mode"])
    gen = dict(synth.CONFIG_B64, target_mode=mode)
    with cf.ThreadPoolExecutor(max_workers=threads) as ex:
        pairs = list(ex.map(lambda s: synth.pair(seed=int(s), **gen), seeds))
    for i, (s, t, _, _) in enumerate(pairs):
        assert (len(s), len(t)) == tuple(fx["n_points"][i]), f"generator drift at seed {seeds[i]}"
    max_pts = max(max(len(s), len(t)) for s, t, _, _ in pairs)
    kw = dict(context_kw or {})
    out = {"pairs": len(seeds), "matched": 0, "rot_matched": 0, "trans_matched": 0,
           "near_ties": 0, "mismatches": []}
    with Context(bandwidth=bandwidth, n_voxels=n_voxels, max_points=max_pts, max_batch=batch,
                 precision=precision, **kw) as ctx:
        ctx.set_graphs(True)
        for a in range(0, len(seeds), batch):
            chunk = pairs[a:a + batch]
            res = ctx.register_pairs([t for _, t, _, _ in chunk], [s for s, _, _, _ in chunk])
            for k, r in enumerate(res):
                i = a + k
                gr = _classify(r["rot_index"], fx, "rot", i)
                gt = _classify(r["trans_index"], fx, "trans", i)
                out["rot_matched"] += gr is None
                out["trans_matched"] += gt is None
                out["matched"] += gr is None and gt is None
                for which, gap, idx in (("rot", gr, r["rot_index"]), ("trans", gt, r["trans_index"])):
                    if gap is None:
                        continue
                    # a translation after another rotation bin is the
                    # correlation of a differently rotated source: it
                    # follows the rotation's mismatch and is not judged
                    follows = which == "trans" and gr is not None
                    tie = (not follows) and gap <= NEAR_TIE
                    out["near_ties"] += tie
                    out["mismatches"].append({"seed": int(seeds[i]), "grid": which,
                                              "gpu_index": int(idx),
                                              "oracle_index": int(fx[f"{which}_index"][i]),
                                              # None: not among the oracle's top entries
                                              "oracle_rel_gap": gap if np.isfinite(gap) else None,
                                              "near_tie": bool(tie),
                                              "follows_rotation": bool(follows)})
    out["rate"] = out["matched"] / max(out["pairs"], 1)
    return out


def fixture_path(root, mode=0):
    return os.path.join(root, "tests", "golden", f"parity_b64_n128_mode{mode}.npz")
