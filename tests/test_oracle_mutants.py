"""Shows that the oracle's pins catch plausible mistakes: each case plants
one mistake in a copy of oracle/phaser_oracle.c, builds it into a temporary
directory, and runs the composition pins against that build in a child
process (PHZ_ORACLE_LIB), which must fail the named test.  The product and
the oracle in the tree are untouched.

Mistakes:
* the SO(3) combine weight inverted, 2 pi sqrt((2l+1)/2) for
  2 pi sqrt(2/(2l+1)) (so3_correlate_fftw.c:120);
* the B^2 coefficient FFTShift / IFFTShift around the spherical fusion
  dropped (spherical-correlation-laplace.cc:23-31, 101-113);
* the N^3 FFTShift / IFFTShift around the spatial fusion dropped
  (spatial-correlation-laplace.cc:36, 87-88);
* with spatial zero padding: the IFFTShift of the padded buffer dropped, the
  padded index map with x and y swapped (spatial-correlation.cc:223-230),
  and the padded size taken as (N + 2 pad)^3 instead of the reference's
  truncated double product (spatial-correlation.cc:17-24);
* the z-score detector's window moved to include the current index, and
  its filtered inputs processed instead of skipped (z-score-peak-
  extraction.cc:77-79,107-110).
"""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "oracle", "phaser_oracle.c")
PINS = os.path.join(ROOT, "tests", "test_oracle_compositions.py")
PEAK_PINS = os.path.join(ROOT, "tests", "test_peak_eval_oracle.py")


def _weight(s):
    return [("const double wigNorm = 2. * M_PI * sqrt(2. / (2. * l + 1.));",
             "const double wigNorm = 2. * M_PI * sqrt((2. * l + 1.) / 2.);")]


def _spherical_shift(s):
    return [("      po_fftshift_real(re, nc); /* spherical-correlation-laplace.cc:101-106 */\n"
             "      po_fftshift_real(im, nc);\n", ""),
            ("  po_ifftshift_real(fr, nc);\n  po_ifftshift_real(fi, nc);\n"
             "  po_ifftshift_real(hr, nc);\n  po_ifftshift_real(hi, nc);\n", "")]


def _spatial_shift(s):
    return [("      po_fftshift(Fc + 2 * (size_t)ch * T, T);\n", ""),
            ("      po_fftshift(Gc + 2 * (size_t)ch * T, T);\n", ""),
            ("  if (shifted) po_ifftshift(C, T);\n", "")]


def _padded_shift(s):
    return [("  if (shifted) po_ifftshift(C, T);\n", "")]


def _pad_map(s):
    return [("      p = sub2ind(ijk[0] + pad, ijk[1] + pad, ijk[2] + pad, Np, Np);",
             "      p = sub2ind(ijk[1] + pad, ijk[0] + pad, ijk[2] + pad, Np, Np);")]


def _zscore_window(s):
    return [("    const uint32_t start = (uint32_t)(i - lag);",
             "    const uint32_t start = (uint32_t)(i - lag) + 1u;")]


def _zscore_filter(s):
    return [("    if (x <= filter_threshold) continue;",
             "    if (x <= filter_threshold) { in[i] = x; }")]


def _padded_size(s):
    return [("  return (uint32_t)(T * factor);",
             "  (void)factor;\n  return (uint32_t)(N + padding_per_dim) * (N + padding_per_dim) *"
             " (N + padding_per_dim);")]


@pytest.mark.parametrize("mutation,caught_by,pins", [
    (_weight, "test_so3_correlation", PINS),
    (_spherical_shift, "test_spherical_fusion_is_shift_fuse_unshift", PINS),
    (_spatial_shift, "test_spatial_fusion_is_shift_fuse_unshift", PINS),
    (_padded_shift, "test_padded_inverse_of_one_product", PINS),
    (_pad_map, "test_padded_inverse_of_one_product", PINS),
    (_padded_size, "test_padded_size_truncates_like_the_reference", PINS),
    (_zscore_window, "test_zscore_window_excludes_the_current_index", PEAK_PINS),
    (_zscore_filter, "test_zscore_filter_skips_without_updating", PEAK_PINS),
])
def test_pins_fail_a_planted_mistake(tmp_path, mutation, caught_by, pins):
    from oracle import build
    src = open(SRC).read()
    for old, new in mutation(src):
        assert src.count(old) == 1, f"mutation site not found: {old!r}"
        src = src.replace(old, new)
    mutant_c = tmp_path / "phaser_oracle_mutant.c"
    mutant_c.write_text(src)
    lib = build.build(src=str(mutant_c), out=str(tmp_path / "libphaser_oracle_mutant.so"))
    env = dict(os.environ, PHZ_ORACLE_LIB=lib)
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-p", "no:cacheprovider",
                        "-p", "no:xdist", pins], cwd=ROOT, env=env, capture_output=True,
                       text=True, timeout=600)
    failed = [l for l in r.stdout.splitlines() if l.startswith("FAILED")]
    assert r.returncode == 1, r.stdout[-2000:] + r.stderr[-2000:]
    assert any(caught_by in l for l in failed), failed
