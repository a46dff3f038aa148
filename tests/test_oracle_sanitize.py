"""The oracle's C restatement built with AddressSanitizer and
UndefinedBehaviorSanitizer (gcc, host code only) and driven through every
option of its registration and the z-score detector on small clouds
(tests/cpp/oracle_sanitize.c): no report, every case finishes."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_oracle_under_address_and_undefined_behaviour_sanitizers(tmp_path):
    exe = tmp_path / "oracle_sanitize"
    r = subprocess.run(["gcc", "-O1", "-g", "-std=gnu11", "-ffp-contract=off",
                        "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
                        "-I", os.path.join(ROOT, "oracle"), "-o", str(exe),
                        os.path.join(ROOT, "tests", "cpp", "oracle_sanitize.c"),
                        os.path.join(ROOT, "oracle", "phaser_oracle.c"), "-lm"],
                       capture_output=True, text=True)
    if r.returncode != 0 and "libasan" in r.stderr:
        pytest.skip("gcc's sanitizer runtimes are absent: " + r.stderr[-300:])
    assert r.returncode == 0, r.stderr
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=600,
                       env=dict(os.environ, ASAN_OPTIONS="detect_leaks=1"))
    print(r.stdout[-2000:], r.stderr[-4000:])
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr
    assert r.returncode == 0 and "all ok" in r.stdout
