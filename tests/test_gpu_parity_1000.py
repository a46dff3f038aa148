"""Argmax-match rate at the metric's configuration (B = 64, 128^3 voxels,
128 x 1024-ray scans, target 70% overlap, noise 2 cm) over seeds 0..999,
against the FP64 oracle's fixtures (tests/golden/make_parity_1000.py): the
rigidly moved target (target_mode 0) and the target re-cast from the moved
sensor (target_mode 1).

Bar: the rotation bin and the translation voxel equal the oracle's on at
least 99.9% of the pairs, and every pair that differs differs by a genuine
near-tie of the oracle's own grid (the device's index is among the oracle's
top entries, within 1e-4 of the grid's largest magnitude of the oracle's
maximum), which FP32 arithmetic cannot order.  Both precision forms.
"""
import os

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.parametrize("mode,precision", [(0, 0), (1, 0), (0, 1)])
def test_argmax_match_1000_pairs(mode, precision):
    from phaser_amd import parity
    path = parity.fixture_path(ROOT, mode)
    if not os.path.exists(path):
        pytest.skip(f"{path} not written (tests/golden/make_parity_1000.py)")
    fx = parity.load(path)
    print(f"parity sweep target_mode={mode} precision={precision} pairs={len(fx['seeds'])}",
          flush=True)
    r = parity.sweep(fx, precision=precision)
    print(f"matched {r['matched']}/{r['pairs']} (rotation {r['rot_matched']}, translation "
          f"{r['trans_matched']}), near-ties {r['near_ties']}", flush=True)
    for m in r["mismatches"]:
        print("  mismatch", m, flush=True)
    assert r["pairs"] == 1000
    assert r["rate"] >= 0.999
    assert all(m["near_tie"] for m in r["mismatches"]), r["mismatches"]
