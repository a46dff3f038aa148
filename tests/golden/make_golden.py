"""Writes oracle fixtures for registrations too slow to recompute inside the
GPU test run (B=128 with 256^3 voxels takes ~80 s per pair in the FP64 oracle).

* oracle_b128_n256.npz: B = 128, 256^3 voxels, ~200k-point scans.
* oracle_b150_pad10_n201_lp85_115.npz: the flags of the reference's tuned
  driver run (phaser_share/run_phaser_core_driver:11-21): bandwidth 150,
  spherical zero padding 10 (SO(3) grid 320^3), 201^3 voxels over +-50 m,
  spatial low-pass bounds 85 / 115 (the S^2 nearest-neighbour acceptance
  radius ((85 + 115) / 201)^2, point-cloud.cc:66-72), on 128 x 1024-ray
  scans with 70% overlap.

For each seed: argmax indices and values, the top-2 values of each grid (for
the near-tie rule), the pose, and the grid values at 4096 fixed sample
indices.  Usage: python tests/golden/make_golden.py [b128] [b150]
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)


def make(bw, N, seeds, cfg, out, **flags):
    """flags: further po.register flags (sph_pad, lp_lower, lp_upper,
    spat_pad); the grids are (2 (bw + sph_pad))^3 and (N + 2 spat_pad)^3."""
    from oracle import oracle as po
    from phaser_amd import synth
    rec = {}
    rng = np.random.default_rng(12345)
    no = 2 * (bw + flags.get("sph_pad", 0))
    Np = N + 2 * flags.get("spat_pad", 0)
    rot_idx = np.sort(rng.choice(no ** 3, 4096, replace=False)).astype(np.int64)
    tr_idx = np.sort(rng.choice(Np ** 3, 4096, replace=False)).astype(np.int64)
    rec["rot_sample_index"] = rot_idx
    rec["trans_sample_index"] = tr_idx
    for s in seeds:
        src, tgt, _, _ = synth.pair(seed=s, **cfg)
        r = po.register(tgt, src, bw=bw, n_voxels=N, want_grids=True, **flags)
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


DRIVER_RUN = dict(bw=150, N=201, sph_pad=10, lp_lower=85, lp_upper=115)


if __name__ == "__main__":
    from phaser_amd import synth
    which = sys.argv[1:] or ["b128", "b150"]
    if "b128" in which:
        make(128, 256, [0, 1], synth.CONFIG_B128, os.path.join(HERE, "oracle_b128_n256.npz"))
    if "b150" in which:
        d = dict(DRIVER_RUN)
        make(d.pop("bw"), d.pop("N"), [0, 1], synth.CONFIG_B64,
             os.path.join(HERE, "oracle_b150_pad10_n201_lp85_115.npz"), **d)
