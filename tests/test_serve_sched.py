"""Scheduling of the request loop (phaser_amd/csrc/host/phaser_serve.cc) on
the CPU: the loop is compiled against a host-only stand-in of the phaser_amd
entry points (tests/cpp/fake_phz.cc) whose registrations take a fixed time
and return indices derived from the clouds, and tests/cpp/serve_sched_test.cc
checks continuous batching across shapes, overlap of contexts, admission by
memory budget with retirement, shared-target batches, rejected submits,
failed context creation, polling and shutdown."""
import os
import subprocess

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
