"""PLY ingestion (include/phaser_ply.h) on the CPU: ASCII and binary files of
both byte orders, every scalar type, list properties and other elements
skipped, field names from the flags' defaults (ply-helper.cc:8-22) or given,
the reference's required fields (x, y, z, intensity: CHECK_GT there, -1
here) and optional ones (reflectivity, ambient, range: zero / absent)."""
import numpy as np
import pytest

from phaser_amd import ply


def _cloud(n, seed=0):
    rng = np.random.default_rng(seed)
    return {
        "x": ("float", rng.normal(size=n).astype(np.float32) * 10),
        "y": ("float", rng.normal(size=n).astype(np.float32) * 10),
        "z": ("float", rng.normal(size=n).astype(np.float32)),
        "intensity": ("float", rng.random(n).astype(np.float32) * 100),
        "reflectivity": ("ushort", rng.integers(0, 65535, n).astype(np.uint16)),
        "noise": ("ushort", rng.integers(0, 4000, n).astype(np.uint16)),
        "range": ("uint", rng.integers(0, 100000, n).astype(np.uint32)),
    }


@pytest.mark.parametrize("fmt", ["ascii", "binary_little_endian", "binary_big_endian"])
def test_read_all_formats(tmp_path, fmt):
    cols = _cloud(257, seed=1)
    p = tmp_path / f"c_{fmt}.ply"
    ply.write_ply(p, cols, fmt=fmt, faces=[[0, 1, 2], [3, 4, 5, 6]])
    c = ply.read_ply(p)
    want = np.stack([cols[k][1].astype(np.float32) for k in ("x", "y", "z", "intensity")], 1)
    np.testing.assert_array_equal(c["xyzi"], want)
    np.testing.assert_array_equal(c["extra"][:, 0], cols["reflectivity"][1].astype(np.float32))
    np.testing.assert_array_equal(c["extra"][:, 1], cols["noise"][1].astype(np.float32))
    np.testing.assert_array_equal(c["range"], cols["range"][1].astype(np.float64))
    assert c["has_reflectivity"] and c["has_ambient"]


def test_scalar_types_and_field_names(tmp_path):
    n = 50
    rng = np.random.default_rng(2)
    cols = {
        "intensity_raw": ("uchar", rng.integers(0, 255, n).astype(np.uint8)),
        "x": ("double", rng.normal(size=n) * 3),
        "y": ("short", rng.integers(-300, 300, n).astype(np.int16)),
        "z": ("char", rng.integers(-100, 100, n).astype(np.int8)),
        "refl": ("int", rng.integers(-5, 5, n).astype(np.int32)),
        "amb": ("double", rng.random(n)),
    }
    for fmt in ("ascii", "binary_big_endian"):
        p = tmp_path / f"t_{fmt}.ply"
        ply.write_ply(p, cols, fmt=fmt)
        c = ply.read_ply(p, intensity="intensity_raw", reflectivity="refl", ambient="amb")
        np.testing.assert_array_equal(c["xyzi"][:, 0], cols["x"][1].astype(np.float32))
        np.testing.assert_array_equal(c["xyzi"][:, 1], cols["y"][1].astype(np.float32))
        np.testing.assert_array_equal(c["xyzi"][:, 2], cols["z"][1].astype(np.float32))
        np.testing.assert_array_equal(c["xyzi"][:, 3], cols["intensity_raw"][1].astype(np.float32))
        np.testing.assert_array_equal(c["extra"][:, 0], cols["refl"][1].astype(np.float32))
        np.testing.assert_array_equal(c["extra"][:, 1], cols["amb"][1].astype(np.float32))
        assert c["range"] is None


def test_optional_fields_missing_give_zero_channels(tmp_path):
    cols = {k: v for k, v in _cloud(20, seed=3).items() if k in ("x", "y", "z", "intensity")}
    p = tmp_path / "xyzi.ply"
    ply.write_ply(p, cols)
    c = ply.read_ply(p)
    assert not c["has_reflectivity"] and not c["has_ambient"] and c["range"] is None
    np.testing.assert_array_equal(c["extra"], np.zeros((20, 2), np.float32))


def test_required_fields_and_malformed_files(tmp_path):
    cols = {k: v for k, v in _cloud(10).items() if k in ("x", "y", "z")}
    p = tmp_path / "noint.ply"
    ply.write_ply(p, cols)
    with pytest.raises(ply.PlyError, match="no intensity field 'intensity'"):
        ply.read_ply(p)
    cols = {k: v for k, v in _cloud(10).items() if k in ("x", "z", "intensity")}
    ply.write_ply(p, cols)
    with pytest.raises(ply.PlyError, match="x/y/z"):
        ply.read_ply(p)
    ply.write_ply(p, _cloud(10))
    data = p.read_bytes()
    (tmp_path / "trunc.ply").write_bytes(data[:-7])
    with pytest.raises(ply.PlyError, match="truncated"):
        ply.read_ply(tmp_path / "trunc.ply")
    (tmp_path / "bad.ply").write_bytes(b"plx\nformat ascii 1.0\nend_header\n")
    with pytest.raises(ply.PlyError, match="magic"):
        ply.read_ply(tmp_path / "bad.ply")
    with pytest.raises(ply.PlyError, match="unable to open"):
        ply.read_ply(tmp_path / "missing.ply")
    huge = (b"ply\nformat binary_little_endian 1.0\nelement vertex 99999999999\n"
            b"property float x\nproperty float y\nproperty float z\nproperty float intensity\n"
            b"end_header\n" + bytes(64))
    (tmp_path / "huge.ply").write_bytes(huge)
    with pytest.raises(ply.PlyError, match="exceeds the file size"):
        ply.read_ply(tmp_path / "huge.ply")


def test_ascii_writer_round_trips(tmp_path):
    rng = np.random.default_rng(5)
    pts = (rng.normal(size=(300, 4)) * [30, 30, 3, 50]).astype(np.float32)
    p = tmp_path / "reg.ply"
    ply.write_ply_xyzi(p, pts)
    assert p.read_text().startswith("ply\nformat ascii 1.0\nelement vertex 300\n")
    np.testing.assert_array_equal(ply.read_ply(p)["xyzi"], pts)
