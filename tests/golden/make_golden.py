"""Writes oracle fixtures for registrations too slow to recompute inside the
GPU test run (B=128 with 256^3 voxels takes ~80 s per pair in the FP64 oracle).

For each seed: argmax indices and values, the top-2 values of each grid (for
the near-tie rule), the pose, and the grid values at 4096 fixed sample
indices.  Usage: python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)


def make(bw, N, seeds, cfg, out):
    from oracle import oracle as po
    from phaser_amd import synth
    rec = {}
    rng = np.random.default_rng(12345)
    rot_idx = np.sort(rng.choice((2 * bw) ** 3, 4096, replace=False)).astype(np.int64)
    tr_idx = np.sort(rng.choice(N ** 3, 4096, replace=False)).astype(np.int64)
    rec["rot_sample_index"] = rot_idx
    rec["trans_sample_index"] = tr_idx
    for s in seeds:
        src, tgt, _, _ = synth.pair(seed=s, **cfg)
        r = po.register(tgt, src, bw=bw, n_voxels=N, want_grids=True)
        rc, tc = r["rot_corr"], r["trans_corr"]
        k = f"s{s}_"
        for key in ("rot_index", "trans_index", "rot_value", "trans_value"):
            rec[k + key] = np.array(r[key])
        for key in ("zyz", "quat", "euler_xyz", "translation"):
            rec[k + key] = np.asarray(r[key])
        rec[k + "T_rot"] = r["T_rot"]
        rec[k + "rot_top2"] = np.sort(rc)[-2:]
        rec[k + "trans_top2"] = np.sort(tc)[-2:]
        rec[k + "rot_max_abs"] = np.array(np.max(np.abs(rc)))
        rec[k + "trans_max_abs"] = np.array(np.max(np.abs(tc)))
        rec[k + "rot_sample"] = rc[rot_idx]
        rec[k + "trans_sample"] = tc[tr_idx]
        rec[k + "n_points"] = np.array([len(src), len(tgt)])
        print(f"seed {s}: rot {r['rot_index']} trans {r['trans_index']}", flush=True)
    np.savez_compressed(out, **rec)


if __name__ == "__main__":
    from phaser_amd import synth
    make(128, 256, [0, 1], synth.CONFIG_B128, os.path.join(HERE, "oracle_b128_n256.npz"))
