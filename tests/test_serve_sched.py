"""Scheduling of the request loop (phaser_amd/csrc/host/phaser_serve.cc) on
the CPU: the loop is compiled against a host-only stand-in of the phaser_amd
entry points (tests/cpp/fake_phz.cc) whose registrations take a fixed time
and return indices derived from the clouds, and tests/cpp/serve_sched_test.cc
checks continuous batching across shapes, overlap of contexts, admission by
memory budget with retirement, shared-target batches, rejected submits,
failed context creation, polling and shutdown."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_serve_scheduling_with_fake_backend(tmp_path):
    exe = tmp_path / "serve_sched_test"
    subprocess.run(["g++", "-O1", "-std=c++17", "-pthread", "-Wall", "-Wno-unused-parameter",
                    "-I", os.path.join(ROOT, "include"), "-o", str(exe),
                    os.path.join(ROOT, "tests", "cpp", "serve_sched_test.cc"),
                    os.path.join(ROOT, "tests", "cpp", "fake_phz.cc"),
                    os.path.join(ROOT, "phaser_amd", "csrc", "host", "phaser_serve.cc")],
                   check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "all ok" in r.stdout


CLANG = "/opt/rocm/llvm/bin/clang++"


@pytest.mark.skipif(not os.path.exists(CLANG), reason="clang++ with ThreadSanitizer is absent")
def test_serve_scheduling_under_thread_sanitizer(tmp_path):
    """The same scheduling test built with clang's ThreadSanitizer must report
    no race.  (GCC 11's TSan reports false positives here: it does not model
    the steady-clock condition-variable waits, wait_until / wait_for, that
    the request loop uses.)"""
    exe = tmp_path / "serve_sched_tsan"
    r = subprocess.run([CLANG, "-O1", "-g", "-std=c++17", "-pthread", "-fsanitize=thread",
                        "-I", os.path.join(ROOT, "include"), "-o", str(exe),
                        os.path.join(ROOT, "tests", "cpp", "serve_sched_test.cc"),
                        os.path.join(ROOT, "tests", "cpp", "fake_phz.cc"),
                        os.path.join(ROOT, "phaser_amd", "csrc", "host", "phaser_serve.cc")],
                       capture_output=True, text=True)
    if r.returncode != 0:
        pytest.skip("clang++ -fsanitize=thread cannot build here: " + r.stderr[-500:])
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300,
                       env=dict(os.environ, TSAN_OPTIONS="halt_on_error=0"))
    print(r.stdout[-2000:], r.stderr[-4000:])
    assert "WARNING: ThreadSanitizer" not in r.stderr
    assert r.returncode == 0 and "all ok" in r.stdout
