"""The S^2 projection's acosf / atan2f (phaser_amd/csrc/libm_f32.hpp, the
device's own code built for the host) against this host's libm, the one the
oracle and the reference call (spherical-projection.cc:19-41): acosf on every
float in [-1, 1] (and the NaN and out-of-domain patterns), atanf on every
float, atan2f on 2^28 random pairs (all finite patterns; metre-scale
exponents; |x| = |y|) and the special values (tests/cpp/libm_f32_check.cc).
Bit-identical results are what make the device's projected points the
oracle's to the bit."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_projection_libm_matches_host_libm_exhaustively(tmp_path):
    exe = tmp_path / "libm_f32_check"
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-pthread",
                    "-Wno-unknown-pragmas", "-o", str(exe),
                    os.path.join(ROOT, "tests", "cpp", "libm_f32_check.cc"), "-lm"], check=True)
    r = subprocess.run([str(exe), "1"], capture_output=True, text=True, timeout=600)
    print(r.stdout[-3000:])
    assert r.returncode == 0
    assert "acosf 0, atanf 0, atan2f 0 mismatches" in r.stdout
    assert "checked 4294967296 bit patterns" in r.stdout
