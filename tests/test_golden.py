"""CPU checks of the committed oracle fixtures (tests/golden/*.npz): the
generator still produces the clouds they were made from, the stored poses
are the oracle's closed forms of the stored argmax indices, and the stored
peak values are the stored maxima."""
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def test_b128_fixture_is_self_consistent(po):
    from phaser_amd import synth
    d = np.load(os.path.join(HERE, "golden", "oracle_b128_n256.npz"))
    for s in (0, 1):
        k = f"s{s}_"
        src, tgt, _, _ = synth.pair(seed=s, **synth.CONFIG_B128)
        assert [len(src), len(tgt)] == list(d[k + "n_points"])
        zyz = po.zyz_from_index(int(d[k + "rot_index"]), 128)
        np.testing.assert_allclose(zyz, d[k + "zyz"], atol=1e-12)
        q = po.quat_from_zyz(zyz)
        np.testing.assert_allclose(q, d[k + "quat"], atol=1e-12)
        np.testing.assert_allclose(po.xyz_from_quat(q), d[k + "euler_xyz"], atol=1e-12)
        xyz = po.ind2sub(int(d[k + "trans_index"]), 256, 256)
        t = [po.translation_from_index(int(v), 256) for v in xyz]
        np.testing.assert_allclose(t, d[k + "translation"], atol=1e-12)
        assert d[k + "rot_top2"][1] == d[k + "rot_value"] == d[k + "rot_max_abs"]
        assert d[k + "trans_top2"][1] == d[k + "trans_value"]
        assert d[k + "rot_top2"][0] <= d[k + "rot_top2"][1]
        assert np.all(d[k + "rot_sample"] <= d[k + "rot_value"])
        assert np.all(d[k + "trans_sample"] <= d[k + "trans_value"])
    assert len(np.unique(d["rot_sample_index"])) == 4096
    assert d["rot_sample_index"].max() < 256 ** 3 and d["trans_sample_index"].max() < 256 ** 3
