"""Builds the in-tree shared libraries of phaser_amd.

* lib/libphaser_amd.so   -- HIP kernels for gfx950 + the C ABI (include/phaser_amd.h)
* lib/libphaser_host.so  -- the C++ mirror of the reference's phaser_core classes
                            (BaseRegistration, SphOptRegistration, ...) over the C ABI
* lib/libphaser_serve.so -- continuous-batching request loop over the C ABI
                            (include/phaser_serve.h)
* lib/libphaser_synth.so -- host-side seeded scan-pair generator

Run `python -m phaser_amd.build` (or __graft_entry__.build()).
"""
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "lib")
OBJ = os.path.join(HERE, "_build")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")
ARCH = os.environ.get("PHZ_OFFLOAD_ARCH", "gfx950")

HIP_SOURCES = ["kernels_s2.hip", "kernels_so3.hip", "kernels_spatial.hip", "kernels_mfma.hip",
               "kernels_peaks.hip", "kernels_pre.hip", "noalloc.hip", "capi.hip"]
CPP_SOURCES = ["tables.cpp"]
HOST_SOURCES = ["host/phaser_host.cc", "host/ply_reader.cc"]

HIP_FLAGS = [f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-Wall",
             "-Wno-unused-function", "-Wno-unused-variable", "-Wno-unused-value",
             "-Wno-unused-result",
             # packed FP32 instructions stay off on gfx950
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-I", os.path.join(ROOT, "include")]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"build step failed ({' '.join(cmd)}):\n{r.stdout}\n{r.stderr}")
    err = "\n".join(l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l and l.strip())
    if verbose and err:
        print(err)


def _compile(src, verbose):
    path = os.path.join(CSRC, src)
    obj = os.path.join(OBJ, os.path.basename(src).replace(".", "_") + ".o")
    if src.endswith(".hip"):
        cmd = [HIPCC, *HIP_FLAGS, "-x", "hip", "-c", path, "-o", obj]
    else:
        cmd = [HIPCC, "-O3", "-std=c++17", "-fPIC", "-Wall", "-I", os.path.join(ROOT, "include"),
               "-c", path, "-o", obj]
    _run(cmd, verbose)
    return obj


def build(verbose=False, jobs=8):
    os.makedirs(LIB, exist_ok=True)
    os.makedirs(OBJ, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(lambda s: _compile(s, verbose), HIP_SOURCES + CPP_SOURCES))
    out = os.path.join(LIB, "libphaser_amd.so")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", out + ".tmp", *objs], verbose)
    os.replace(out + ".tmp", out)
    # host C++ mirror of the reference classes, linked against the C ABI
    host_src = [os.path.join(CSRC, s) for s in HOST_SOURCES if os.path.exists(os.path.join(CSRC, s))]
    if host_src:
        hout = os.path.join(LIB, "libphaser_host.so")
        _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall", "-ffp-contract=off",
              "-I", os.path.join(ROOT, "include"), "-o", hout + ".tmp", *host_src,
              "-L", LIB, "-lphaser_amd", "-Wl,-rpath,$ORIGIN"], verbose)
        os.replace(hout + ".tmp", hout)
        drv = os.path.join(ROOT, "tests", "cpp", "host_driver.cc")
        if os.path.exists(drv):
            dout = os.path.join(LIB, "phaser_host_driver")
            _run(["g++", "-O2", "-std=c++17", "-Wall", "-I", os.path.join(ROOT, "include"),
                  "-o", dout + ".tmp", drv, "-L", LIB, "-lphaser_host", "-lphaser_amd",
                  "-Wl,-rpath,$ORIGIN"], verbose)
            os.replace(dout + ".tmp", dout)
        # the reference's command-line driver (driver.cc) over the mirror
        cout = os.path.join(LIB, "phaser_core_driver")
        _run(["g++", "-O2", "-std=c++17", "-Wall", "-I", os.path.join(ROOT, "include"),
              "-o", cout + ".tmp", os.path.join(CSRC, "host", "phaser_core_driver.cc"), "-L", LIB,
              "-lphaser_host", "-lphaser_amd", "-Wl,-rpath,$ORIGIN"], verbose)
        os.replace(cout + ".tmp", cout)
    vout = os.path.join(LIB, "libphaser_serve.so")
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall", "-pthread",
          "-I", os.path.join(ROOT, "include"), "-o", vout + ".tmp",
          os.path.join(CSRC, "host", "phaser_serve.cc"), "-L", LIB, "-lphaser_amd",
          "-Wl,-rpath,$ORIGIN"], verbose)
    os.replace(vout + ".tmp", vout)
    sout = os.path.join(LIB, "libphaser_synth.so")
    _run(["gcc", "-O2", "-std=gnu11", "-fPIC", "-shared", "-Wall", "-o", sout + ".tmp",
          os.path.join(CSRC, "synth.c"), "-lm"], verbose)
    os.replace(sout + ".tmp", sout)
    return out


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
