"""CPU-side checks of the drop-in boundary: the HIP library loads and exports
every entry point declared in include/*.h (no compute calls: no GPU here),
and the ctypes structures match the C layout."""
import ctypes as C
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(phz_[a-z0-9_]+)\s*\(", text)))


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


def test_synth_library_exports():
    from phaser_amd import synth
    L = synth._load()
    for n in _declared("phaser_synth.h"):
        assert hasattr(L, n), n


def test_struct_layouts():
    from phaser_amd import capi
    assert C.sizeof(capi.Config) == 4 * 8 + 8 + 4 + 4
    # rot/trans index + values (16) + zyz,quat,euler (80) + T (64) + translation (24)
    assert C.sizeof(capi.Result) == 16 + 80 + 64 + 24


def test_default_config_and_error_path_without_gpu():
    from phaser_amd import capi
    cfg = capi.default_config()
    assert cfg.bandwidth == 64 and cfg.n_voxels == 128
    assert cfg.discretize_lower == -50 and cfg.discretize_upper == 50
    bad = capi.default_config(bandwidth=7)  # 14 = 2*7 is not a 2/3/5 size
    h = C.c_void_p()
    rc = capi.lib().phz_create(C.byref(bad), C.byref(h))
    assert rc == -1
    assert b"factor" in capi.lib().phz_last_error()
    bad2 = capi.default_config(spherical_zero_padding=3)
    assert capi.lib().phz_create(C.byref(bad2), C.byref(h)) == -1
