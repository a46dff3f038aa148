"""Registration throughput of the MI355X correlation core.

Workload: seeded synthetic scan pairs with 128 rings x 1024 azimuths (about
80k source points), the target rigidly moved by a uniform random rotation and
a translation in +-5 m, cropped to 70% of its azimuth ring, noise 2 cm;
B = 64 (128 x 128 S^2 grid, 128^3 SO(3) bins), 128^3 voxels over +-50 m.
One step registers `--batch` independent pairs whose clouds are already in
HBM (one phz_register_pairs_device call, correlation grids written out), so
value = batch * steps / time.  Latency p50/p99 comes from single-pair calls.
The argmax-match rate compares the rotation bin and translation voxel of the
first --cpu-sample pairs with the FP64 oracle, which is also timed on one
host core as the CPU baseline.

Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
B, N = 64, 128
WORKLOAD = dict(rings=128, azimuths=1024, noise_sigma=0.02, overlap=0.7)


def stage_bytes(B, N, P_mean, pairs):
    """Algorithmic (compulsory) HBM bytes of one launch of each stage for a
    batch of `pairs` independent registrations: inputs read once, outputs
    written once, FP32 / complex64 elements, C = 2 populated channels."""
    n = 2 * B
    C = 2
    H = N * N * (N // 2 + 1)
    clouds = 2 * pairs
    return {
        "s2_cells": clouds * (16 * P_mean + 24 * P_mean),
        "s2_nn": clouds * (24 * P_mean + 16 * n * n + 4 * C * n * n),
        "s2_transform": clouds * (4 * C * n * n + 8 * C * B * B),
        "so3_synth": 4 * n * B * (B + 1) * (B + 2) // 6 + pairs * 8 * n * B * (2 * B - 1),
        "so3_fft_alpha": pairs * (8 * n * B * (2 * B - 1) + 8 * n * B * n),
        "so3_fft_gamma": pairs * (8 * n * B * n + 4 * n ** 3),
        "voxel_scatter": clouds * (16 * P_mean + 4 * (C + 1) * N ** 3),
        "spatial_forward": clouds * C * (4 * N ** 3 + 8 * H),
        "spatial_inverse": pairs * (8 * H * C * 2 + 4 * N ** 3),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--batch", type=int, default=16)
    ap.add_argument("--latency-samples", type=int, default=40)
    ap.add_argument("--cpu-sample", type=int, default=2)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); registrations are "
              "independent, so N GPUs would each run this same benchmark on their own pairs.",
              file=sys.stderr)
        sys.exit(2)

    from phaser_amd import synth
    from phaser_amd.capi import Context

    P = args.batch
    pairs = [synth.pair(seed=s, **WORKLOAD) for s in range(P)]
    max_pts = max(max(len(s), len(t)) for s, t, _, _ in pairs)
    P_mean = float(np.mean([len(s) + len(t) for s, t, _, _ in pairs]) / 2)
    ctx = Context(bandwidth=B, n_voxels=N, max_points=max_pts, max_batch=P)
    d_src = [ctx.device_array(s) for s, _, _, _ in pairs]
    d_tgt = [ctx.device_array(t) for _, t, _, _ in pairs]
    Ps = [len(s) for s, _, _, _ in pairs]
    Pt = [len(t) for _, t, _, _ in pairs]
    d_rot = ctx.device_empty(P * (2 * B) ** 3 * 4)
    d_trans = ctx.device_empty(P * N ** 3 * 4)

    def step():
        return ctx.register_pairs_device([a.ptr for a in d_tgt], Pt, [a.ptr for a in d_src], Ps,
                                         d_rot=d_rot.ptr, d_trans=d_trans.ptr)

    for _ in range(args.warmup):
        step()
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        res = step()
    ctx.synchronize()
    t1 = time.perf_counter()
    elapsed = t1 - t0
    value = P * args.steps / elapsed

    # single-pair latency (p50 / p99)
    lat = []
    for i in range(args.latency_samples):
        k = i % P
        ta = time.perf_counter()
        ctx.register_device(d_tgt[k].ptr, Pt[k], d_src[k].ptr, Ps[k])
        lat.append((time.perf_counter() - ta) * 1e3)
    lat = np.array(lat)

    # per-stage device time over `steps` more batched steps (HIP events on the
    # context stream), dominant stage -> roofline
    ctx.set_profiling(True)
    for _ in range(args.steps):
        step()
    st = ctx.stage_times()
    ctx.set_profiling(False)
    alg = stage_bytes(B, N, P_mean, P)
    dom = max((k for k in alg if st.get(k, (0, 0))[1] > 0), key=lambda k: st[k][0])
    dom_ms = st[dom][0] / st[dom][1]
    achieved = alg[dom] / (dom_ms * 1e-3) / 1e9
    stage_ms = {k: round(v[0] / max(v[1], 1), 4) for k, v in st.items()}

    # oracle: argmax match and CPU baseline on the first pairs
    from oracle import oracle as po
    S = min(args.cpu_sample, P)
    match = 0
    tc0 = time.perf_counter()
    for k in range(S):
        s, t, _, _ = pairs[k]
        o = po.register(t, s, bw=B, n_voxels=N, kdtree=True)
        match += int(o["rot_index"] == res[k]["rot_index"] and o["trans_index"] == res[k]["trans_index"])
    cpu_s = time.perf_counter() - tc0

    out = {
        "metric": "registrations/sec at B=64 + 128^3 voxels",
        "value": round(value, 2),
        "unit": "registrations/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic (seeded ray-cast scenes, seeds 0..batch-1)",
        "config": {"workload": "B=64 S2 grid 128x128, SO(3) 128^3, voxels 128^3 over +-50 m, "
                               "128x1024-ray scans, target 70% overlap, noise 2 cm",
                   "pairs_per_step": P, "mean_points_per_cloud": round(P_mean),
                   "form": "plain fp32"},
        "p50_ms": round(float(np.percentile(lat, 50)), 4),
        "p99_ms": round(float(np.percentile(lat, 99)), 4),
        "argmax_match": {"matched": match, "pairs": S, "rate": round(match / max(S, 1), 4)},
        "stage_ms_per_step": stage_ms,
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 2),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5),
                     "traffic": None, "bytes_per_launch": int(alg[dom]),
                     "avg_launch_ms": round(dom_ms, 4)},
        "cpu_baseline": {"value": round(S / cpu_s, 4), "unit": "registrations/s", "cores": 1,
                         "kind": "port",
                         "sample": f"{S} pairs of the same workload through the FP64 oracle "
                                   f"(kd-tree NN, own FFT), {cpu_s:.1f} s"},
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
