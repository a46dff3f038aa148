"""Registration throughput of the MI355X correlation core.

Workload: seeded synthetic scan pairs with 128 rings x 1024 azimuths (about
80k source points), the target cast again from the sensor moved by a uniform
random rotation and a translation in +-5 m (--target-mode 1, the default; 0
moves the source's points rigidly instead), cropped to 70% of its azimuth
ring, noise 2 cm;
B = 64 (128 x 128 S^2 grid, 128^3 SO(3) bins), 128^3 voxels over +-50 m.
One step registers `--batch` independent pairs whose clouds are already in
HBM, split over `--streams` contexts (one phz_register_pairs_device call per
context, both correlation grids of every pair written to HBM), so value =
batch * steps / time.  Latency p50/p99 comes from single-pair calls.  The
argmax-match rate registers the 1000 seeded pairs of the committed FP64
oracle fixture (tests/golden/parity_b64_n128_mode{0,1}.npz) and compares the
rotation bin and translation voxel of each (the fixture of the same target
mode; the other mode's fixture is swept too, FP32 form); the FP64 oracle itself is timed
on --cpu-sample pairs over up to --cpu-threads host threads (one pair per
thread) as the CPU baseline.  The roofline names the kernel of the stage
with the largest device time: its algorithmic bytes per launch over its mean
launch duration (HIP events on the context stream), against the 8 TB/s HBM
peak; the streaming copy kernel's rate is reported beside it.  `served`:
a short Poisson sweep of the serving mix through the request loop, per
precision form.

Prints one JSON line.
"""
import argparse
import concurrent.futures as cf
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


def kernel_bytes(B, N, P_mean, pairs, shared=False, plane=False, fused_reg=False,
                 lines_fwd=False, lines_inv=False, nn_lanes=4, occ=False, inv_plane=False,
                 so3_reg=False):
    """Algorithmic (compulsory) HBM bytes of ONE launch of the kernel that
    dominates each stage, for a batch of `pairs` registrations: inputs read
    once, outputs written once, FP32 / complex64 elements, C = 2 populated
    channels (range/intensity), H = N^2 (N/2+1) half-spectrum entries.
    Returns {stage: (kernel name, bytes per launch)}.  The names follow the
    tuned pipeline: with occ (occupancy passes, fused register kernels) the
    forward x pass is k_fft_x_occ, the r2c k_spat_r2c_rows and the cross
    spectrum k_spat_cross_zz_reg, which read only the voxel rows / z planes
    that hold points, so their measured traffic is below these dense bytes;
    with inv_plane the inverse x + c2r is k_spat_inv_plane."""
    n = 2 * B
    C = 2
    H = N * N * (N // 2 + 1)
    nt = 1 if shared else pairs
    clouds = nt + pairs
    out = {
        # S^2 sampling per cloud: its points read once (16 B each) and its C
        # grids written once (the query grid and the kernel's own sorted
        # layout are not counted)
        "s2_nn": (f"k_nn<{nn_lanes}>", clouds * (16 * P_mean + 4 * C * n * n)),
        # Wigner table once per batch + fused coefficients in, S half-plane out
        "so3_synth": ("k_so3_synth", 4 * n * B * (B + 1) * (B + 2) // 6
                      + clouds * 8 * B * B + pairs * 8 * n * B * (2 * B - 1)),
        "so3_fft_alpha": ("k_so3_fft_alpha", pairs * (8 * n * B * (2 * B - 1) + 8 * n * B * n)),
        "so3_fft_gamma": ("k_so3_fft_gamma", pairs * (8 * n * B * n + 4 * n ** 3)),
        # points in; three FP32 accumulators read-modify-written per point
        "voxel_scatter": ("k_voxel_scatter<false>", clouds * (16 + 24) * P_mean),
        # three N^3 accumulators in, two half spectra out (the plane form also
        # transforms along x on the way)
        "spatial_r2c": ("k_spat_plane_fwd" if plane else
                        ("k_spat_r2c_reg" if fused_reg else "k_spat_r2c"),
                        clouds * (12 * N ** 3 + 2 * 8 * H)),
        # one FFT pass (x or z) over every half spectrum: read + write
        "spatial_fft_fwd": ("k_fft_lines<-1>" if lines_fwd else "k_fft_cols<-1>",
                            2 * clouds * 16 * H),
        # C target + C source spectra in, cross spectrum out
        "spatial_cross_z": ("k_spat_cross_z_reg" if fused_reg else "k_spat_cross_z",
                            nt * C * 8 * H + pairs * (C * 8 * H + 8 * H)),
        "spatial_fft_inv": ("k_fft_lines<1>" if lines_inv else "k_fft_cols<1>", pairs * 16 * H),
        "spatial_c2r": ("k_spat_c2r_reg" if fused_reg else "k_spat_c2r",
                        pairs * (8 * H + 4 * N ** 3)),
    }
    rename = {}
    if so3_reg:
        rename.update(so3_fft_alpha="k_so3_fft_alpha_reg", so3_fft_gamma="k_so3_fft_gamma_reg")
    if occ and fused_reg:
        rename.update(spatial_r2c="k_spat_r2c_rows", spatial_fft_fwd="k_fft_x_occ",
                      spatial_cross_z="k_spat_cross_zz_reg")
    if inv_plane and fused_reg:
        rename.update(spatial_c2r="k_spat_inv_plane")
    for stage, name in rename.items():
        out[stage] = (name, out[stage][1])
    return out


FORM_NAMES = {0: "plain fp32", 1: "split fp16 mfma contractions"}


def host_cpu_model():
    """The host CPU's model name (/proc/cpuinfo), or None."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None

# launches of the stage's dominant kernel per pipeline pass
LAUNCHES = {"spatial_fft_fwd": 2, "voxel_scatter": 2}


def concurrent_throughput(args, pairs, max_pts, tuning, grids=True):
    """The batch split over args.streams contexts (each with its own HIP
    stream and buffers), driven from one host thread each so that their
    kernels overlap on the GPU; returns (registrations/s, seconds) over
    args.steps steps of every context, timed from a common start.  grids:
    every pair's rotation and translation grids are written to HBM (the
    reference returns the rotation correlation with its result,
    sph-opt-registration.cc:71)."""
    import threading
    from phaser_amd.capi import Context
    S, P = args.streams, len(pairs)
    parts = [list(range(k, P, S)) for k in range(S)]
    t = tuning
    subs = []
    for idx in parts:
        c = Context(bandwidth=B, n_voxels=N, max_points=max_pts, max_batch=len(idx),
                    precision=args.precision)
        c.set_tuning(t["fwd_x"], t["fwd_z"], t["inv_x"], t["plane_fwd"], t["nn_lanes"],
                     t["fused_reg"], t["so3_reg"], t["occ_fwd"], t["inv_plane"], t["nn_xcd"],
                     t["nn_seed"])
        if args.graphs:
            c.set_graphs(True)
        ds = [c.device_array(pairs[i][0]) for i in idx]
        dt = [c.device_array(pairs[i][1]) for i in idx]
        outs = ([c.device_empty(len(idx) * (2 * B) ** 3 * 4), c.device_empty(len(idx) * N ** 3 * 4)]
                if grids else [])
        subs.append((c, ds, dt, [len(pairs[i][0]) for i in idx], [len(pairs[i][1]) for i in idx],
                     outs))
    start = threading.Barrier(S + 1)
    done = threading.Barrier(S + 1)
    errors = []

    def run(c, ds, dt, ps, pt, outs):
        kw = dict(d_rot=outs[0].ptr, d_trans=outs[1].ptr) if outs else {}
        try:
            for _ in range(args.warmup):
                c.register_pairs_device([a.ptr for a in dt], pt, [a.ptr for a in ds], ps, **kw)
            c.synchronize()
            start.wait()
            for _ in range(args.steps):
                c.register_pairs_device([a.ptr for a in dt], pt, [a.ptr for a in ds], ps, **kw)
            c.synchronize()
        except Exception as e:  # reported after the join
            errors.append(e)
            start.abort()
        finally:
            done.wait()

    threads = [threading.Thread(target=run, args=sub) for sub in subs]
    for th in threads:
        th.start()
    start.wait()
    t0 = time.perf_counter()
    done.wait()
    elapsed = time.perf_counter() - t0
    for th in threads:
        th.join()
    for c, ds, dt, _, _, outs in subs:
        for a in ds + dt + outs:
            a.free()
        c.close()
    if errors:
        raise errors[0]
    return P * args.steps / elapsed, elapsed


def variant_throughput(args, pairs, max_pts, tuning, steps=5, warmup=2):
    """Registrations/s of the rows beyond the plain path on the same pairs,
    one context and stream each (HIP graphs, the main run's tuning), against
    the plain path measured the same way: clouds with reflectivity and ambient
    (four live channels), peak-window fits with 3 neighbours each,
    spherical zero padding 16 (SO(3) grid of 160^3) and SphRegistration's
    blended single channel (method 1)."""
    from phaser_amd.capi import Context
    P = len(pairs)
    B, N = 64, 128
    Ps = [len(s) for s, _, _, _ in pairs]
    Pt = [len(t) for _, t, _, _ in pairs]

    def extras(cloud):
        i = cloud[:, 3].astype(np.float32)
        return np.stack([i * i / np.float32(100), np.float32(50) + np.float32(0.5) * i], 1)

    def run(name, rot_side, **kw):
        ctx = Context(bandwidth=B, n_voxels=N, max_points=max_pts, max_batch=P, **kw)
        t = tuning
        ctx.set_tuning(t["fwd_x"], t["fwd_z"], t["inv_x"], t["plane_fwd"], t["nn_lanes"],
                       t["fused_reg"], t["so3_reg"], t["occ_fwd"], t["inv_plane"], t["nn_xcd"],
                     t["nn_seed"])
        if args.graphs:
            ctx.set_graphs(True)
        ds = [ctx.device_array(s) for s, _, _, _ in pairs]
        dt = [ctx.device_array(t) for _, t, _, _ in pairs]
        four = kw.get("channels") == 4
        dse = [ctx.device_array(extras(s)) for s, _, _, _ in pairs] if four else []
        dte = [ctx.device_array(extras(t)) for _, t, _, _ in pairs] if four else []
        d_rot = ctx.device_empty(P * rot_side ** 3 * 4)
        d_trans = ctx.device_empty(P * N ** 3 * 4)

        def step():
            if four:
                return ctx.register_pairs_channels_device(
                    [a.ptr for a in dt], [a.ptr for a in dte], Pt, [a.ptr for a in ds],
                    [a.ptr for a in dse], Ps, d_rot=d_rot.ptr, d_trans=d_trans.ptr)
            return ctx.register_pairs_device([a.ptr for a in dt], Pt, [a.ptr for a in ds], Ps,
                                             d_rot=d_rot.ptr, d_trans=d_trans.ptr)
        for _ in range(warmup):
            step()
        ctx.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            step()
        ctx.synchronize()
        dt_s = time.perf_counter() - t0
        for a in ds + dt + dse + dte + [d_rot, d_trans]:
            a.free()
        ctx.close()
        return {"value": round(P * steps / dt_s, 2), "ms_per_step": round(dt_s / steps * 1e3, 4)}

    out = {"plain_one_stream": run("plain", 2 * B)}
    out["four_channels"] = run("four", 2 * B, channels=4)
    out["peak_neighbors_3"] = run("peak", 2 * B, bingham_peak_neighbors=3,
                                  gaussian_peak_neighbors=3)
    out["spherical_zero_padding_16"] = run("pad", 2 * (B + 16), spherical_zero_padding=16)
    out["sph_method"] = run("sph", 2 * B, method=1)
    return out


def load_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 --pmc
    summary (FETCH_SIZE + WRITE_SIZE, scaled by the same counters on the
    streaming copy kernel), or None when no summary covers it."""
    from phaser_amd import profiling
    path = os.path.join(ROOT, "profiles", "traffic_b64_n128_batch64.json")
    try:
        with open(path) as f:
            return profiling.lookup(json.load(f), kernel)
    except (OSError, ValueError):
        return None


SERVE_SHAPES = [(32, 64), (64, 128), (96, 192), (128, 256)]


def serve_pool(args):
    from phaser_amd import synth
    with cf.ThreadPoolExecutor(max_workers=16) as ex:
        return {sh: list(ex.map(lambda seed: synth.mixed_pair(seed, sh[0]), range(args.pool)))
                for sh in SERVE_SHAPES}


def serve_sweep(args, pool, precision, rates, duration):
    """Poisson arrivals of scan pairs with bandwidth B in {32, 64, 96, 128}
    (voxel grids 64^3 .. 256^3, overlap 40-90%, noise 1-5 cm, 5% outliers)
    through the continuous-batching request loop, at each offered rate:
    throughput and p50/p99 latency (submit -> result).  Returns (sweep, server
    statistics)."""
    from phaser_amd import serve
    rng = np.random.default_rng(0)
    max_pts = max(max(len(s), len(t)) for v in pool.values() for s, t, _, _ in v)
    sweep = []
    with serve.Server(max_batch=args.serve_batch, contexts_per_shape=2,
                      hbm_budget=int(args.hbm_budget_gib * (1 << 30)),
                      max_points=max_pts, autotune=bool(args.autotune),
                      precision=precision) as srv:
        for sh in SERVE_SHAPES:  # build the contexts before timing
            src, tgt, _, _ = pool[sh][0]
            srv.wait(srv.submit(tgt, src, bandwidth=sh[0], n_voxels=sh[1]))
        for lam in rates:
            n_req = max(int(lam * duration), 8)
            tickets, shapes = [], []
            t_start = time.perf_counter()
            next_t = t_start
            for _ in range(n_req):
                next_t += rng.exponential(1.0 / lam)
                delay = next_t - time.perf_counter()
                if delay > 0:
                    time.sleep(delay)
                sh = SERVE_SHAPES[int(rng.integers(len(SERVE_SHAPES)))]
                src, tgt, _, _ = pool[sh][int(rng.integers(args.pool))]
                tickets.append(srv.submit(tgt, src, bandwidth=sh[0], n_voxels=sh[1]))
                shapes.append(sh[0])
            got = [srv.wait(t) for t in tickets]
            t_end = time.perf_counter()
            lat = np.array([g["latency_ms"] for g in got])
            per_b = {}
            for b, g in zip(shapes, got):
                per_b.setdefault(b, []).append(g["latency_ms"])
            sweep.append({
                "rate_offered": lam, "requests": n_req,
                "throughput": round(n_req / (t_end - t_start), 2),
                "p50_ms": round(float(np.percentile(lat, 50)), 3),
                "p99_ms": round(float(np.percentile(lat, 99)), 3),
                "mean_batch": round(float(np.mean([g["batch_pairs"] for g in got])), 2),
                "failed": int(sum(g["status"] != 0 for g in got)),
                "p50_ms_by_bandwidth": {str(b): round(float(np.percentile(v, 50)), 3)
                                        for b, v in sorted(per_b.items())}})
            print(f"serve {FORM_NAMES[precision]} rate {lam}: {sweep[-1]}", file=sys.stderr,
                  flush=True)
        st = srv.stats()
    return sweep, {k: int(v) for k, v in st.items()}


def serve_summary(sweep):
    best = max(sweep, key=lambda r: r["throughput"])
    return {"value": best["throughput"], "p50_ms": best["p50_ms"], "p99_ms": best["p99_ms"],
            "at_rate": best["rate_offered"], "sweep": sweep}


def serve_bench(args):
    pool = serve_pool(args)
    rates = [float(x) for x in args.rates.split(",")]
    sweep, st = serve_sweep(args, pool, args.precision, rates, args.duration)
    best = max(sweep, key=lambda r: r["throughput"])
    out = {
        "metric": "served registrations/sec, mixed B in {32,64,96,128}, Poisson arrivals",
        "value": best["throughput"], "unit": "registrations/s", "n_gpus": 1,
        "steps": len(rates), "warmup": len(SERVE_SHAPES), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "fp32", "data": "synthetic",
        "config": {"workload": "serving mix: B 32/64/96/128 with 64^3/128^3/192^3/256^3 "
                               "voxels, overlap 40-90%, noise 1-5 cm, 5% outliers",
                   "max_batch": args.serve_batch, "contexts_per_shape": 2,
                   "hbm_budget_gib": args.hbm_budget_gib, "form": FORM_NAMES[args.precision]},
        "p50_ms": best["p50_ms"], "p99_ms": best["p99_ms"], "sweep": sweep,
        "server": st,
    }
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--batch", type=int, default=64)
    ap.add_argument("--latency-samples", type=int, default=40)
    ap.add_argument("--cpu-sample", type=int, default=16)
    ap.add_argument("--cpu-threads", type=int, default=16)
    ap.add_argument("--mode", choices=["batch", "serve"], default="batch")
    ap.add_argument("--autotune", type=int, default=1, help="pick FFT column tiles first")
    ap.add_argument("--graphs", type=int, default=1, help="replay the pipeline from HIP graphs")
    ap.add_argument("--nn-lanes", type=int, default=0, choices=[0, 1, 2, 4, 8, 16],
                    help="lanes per query of the S2 nearest-neighbour search (0: default)")
    ap.add_argument("--streams", type=int, default=2,
                    help="split each batch over this many contexts/streams driven concurrently")
    ap.add_argument("--nn-xcd", type=int, default=0, choices=[-1, 0, 1],
                    help="XCD-ordered NN workgroups: 1 on, -1 off, 0 library default")
    ap.add_argument("--nn-seed", type=int, default=0, choices=[-1, 0, 1, 2, 3],
                    help="NN search: bit 0 seeds, bit 1 LDS band skipping; -1 neither, "
                         "0 library default")
    ap.add_argument("--inv-plane", type=int, default=0, choices=[-1, 0, 1],
                    help="inverse x + c2r as one pass over z planes in LDS (0: default, on)")
    ap.add_argument("--occ-fwd", type=int, default=0, choices=[-1, 0, 1],
                    help="forward spatial passes skip empty voxel rows/planes (0: default, on)")
    ap.add_argument("--precision", type=int, default=0, choices=[0, 1],
                    help="contractions: 0 fp32 (plain form), 1 split fp16 on MFMA")
    ap.add_argument("--compare-forms", type=int, default=1,
                    help="also time the other precision form on the same pairs")
    ap.add_argument("--variants", type=int, default=1,
                    help="also time four channels, peak windows and spherical padding")
    ap.add_argument("--rates", default="100,200,400,800", help="serve: offered requests/s")
    ap.add_argument("--duration", type=float, default=3.0, help="serve: seconds per rate")
    ap.add_argument("--pool", type=int, default=6, help="serve: pairs per bandwidth")
    ap.add_argument("--serve-batch", type=int, default=8)
    ap.add_argument("--hbm-budget-gib", type=float, default=64.0)
    ap.add_argument("--parity-pairs", type=int, default=1000,
                    help="argmax match over this many fixture pairs (0: the timed CPU sample)")
    ap.add_argument("--serve-rates", default="500,2000",
                    help="batch mode: offered rates of the short served sweep per form ('' = off)")
    ap.add_argument("--serve-duration", type=float, default=2.0)
    ap.add_argument("--target-mode", type=int, default=1, choices=[0, 1],
                    help="1: target re-cast from the moved sensor; 0: source points moved")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); registrations are "
              "independent, so N GPUs would each run this same benchmark on their own pairs.",
              file=sys.stderr)
        sys.exit(2)

    if args.mode == "serve":
        return serve_bench(args)

    from phaser_amd import synth
    from phaser_amd.capi import Context

    P = args.batch
    pairs = [synth.pair(seed=s, target_mode=args.target_mode, **WORKLOAD) for s in range(P)]
    max_pts = max(max(len(s), len(t)) for s, t, _, _ in pairs)
    P_mean = float(np.mean([len(s) + len(t) for s, t, _, _ in pairs]) / 2)
    ctx = Context(bandwidth=B, n_voxels=N, max_points=max_pts, max_batch=P,
                  precision=args.precision)
    tuning = ctx.autotune(repeats=5) if args.autotune else ctx.tuning()
    if args.graphs:
        ctx.set_graphs(True)
    if args.nn_lanes or args.occ_fwd or args.inv_plane or args.nn_xcd or args.nn_seed:
        t = ctx.tuning()
        ctx.set_tuning(t["fwd_x"], t["fwd_z"], t["inv_x"], t["plane_fwd"],
                       args.nn_lanes or t["nn_lanes"], t["fused_reg"], t["so3_reg"],
                       args.occ_fwd or t["occ_fwd"], args.inv_plane or t["inv_plane"],
                       args.nn_xcd or t["nn_xcd"], args.nn_seed or t["nn_seed"])
        tuning = ctx.tuning()
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
    one_stream = {"value": round(value, 2), "ms_per_step": round(elapsed / args.steps * 1e3, 4)}
    no_grids = None
    if args.streams > 1:
        value, elapsed = concurrent_throughput(args, pairs, max_pts, tuning, grids=True)
        v2, e2 = concurrent_throughput(args, pairs, max_pts, tuning, grids=False)
        no_grids = {"value": round(v2, 2), "ms_per_step": round(e2 / args.steps * 1e3, 4)}

    # single-pair latency (p50 / p99)
    lat = []
    for i in range(args.latency_samples):
        k = i % P
        ta = time.perf_counter()
        ctx.register_device(d_tgt[k].ptr, Pt[k], d_src[k].ptr, Ps[k])
        lat.append((time.perf_counter() - ta) * 1e3)
    lat = np.array(lat)

    # per-stage device time over `steps` more batched steps (HIP events on the
    # context stream); the stage with the largest time is the roofline kernel
    ctx.set_profiling(True)
    for _ in range(args.steps):
        step()
    st = ctx.stage_times()
    ctx.set_profiling(False)
    plane = tuning["plane_fwd"] == 1
    # lanes per NN query: the tuned value, or the library's choice by clouds
    # per launch (2 for 64 clouds or more, else 4)
    lanes = tuning["nn_lanes"] or (2 if 2 * P >= 64 else 4)
    kb = kernel_bytes(B, N, P_mean, P, plane=plane, fused_reg=tuning["fused_reg"] == 1,
                      lines_fwd=tuning["fwd_z"] == 1, lines_inv=tuning["inv_x"] == 1,
                      nn_lanes=lanes, occ=tuning["occ_fwd"] == 1 and not plane,
                      inv_plane=tuning["inv_plane"] == 1, so3_reg=tuning["so3_reg"] == 1)
    dom = max((k for k in kb if st.get(k, (0, 0))[1] > 0), key=lambda k: st[k][0])
    per_call = dict(LAUNCHES, spatial_fft_fwd=1 if plane else 2)
    launches = st[dom][1] * per_call.get(dom, 1)
    dom_ms = st[dom][0] / launches
    kname, kbytes = kb[dom]
    achieved = kbytes / (dom_ms * 1e-3) / 1e9
    stage_ms = {k: round(v[0] / max(v[1], 1), 4) for k, v in st.items()}

    # the other precision form on the same pairs (its own context)
    other = None
    if args.compare_forms:
        ctx2 = Context(bandwidth=B, n_voxels=N, max_points=max_pts, max_batch=P,
                       precision=1 - args.precision)
        if args.autotune:
            ctx2.autotune(repeats=5)
        if args.graphs:
            ctx2.set_graphs(True)
        d_src2 = [ctx2.device_array(s) for s, _, _, _ in pairs]
        d_tgt2 = [ctx2.device_array(t) for _, t, _, _ in pairs]
        d_rot2 = ctx2.device_empty(P * (2 * B) ** 3 * 4)
        d_trans2 = ctx2.device_empty(P * N ** 3 * 4)

        def step2():
            return ctx2.register_pairs_device([a.ptr for a in d_tgt2], Pt, [a.ptr for a in d_src2],
                                              Ps, d_rot=d_rot2.ptr, d_trans=d_trans2.ptr)

        for _ in range(args.warmup):
            step2()
        ctx2.synchronize()
        ta = time.perf_counter()
        for _ in range(args.steps):
            res2 = step2()
        ctx2.synchronize()
        el2 = time.perf_counter() - ta
        ctx2.set_profiling(True)
        step2()
        st2 = ctx2.stage_times()
        ctx2.set_profiling(False)
        other = {"precision": 1 - args.precision, "value": P * args.steps / el2,
                 "ms_per_step": el2 / args.steps * 1e3, "res": res2,
                 "stage_ms": {k: round(st2[k][0] / max(st2[k][1], 1), 4)
                              for k in ("s2_transform", "so3_synth")}}
        for a in d_src2 + d_tgt2 + [d_rot2, d_trans2]:
            a.free()
        ctx2.close()

    # HBM reference: the streaming copy kernel on 1 GiB
    nb = 1 << 30
    d_a, d_b = ctx.device_empty(nb), ctx.device_empty(nb)
    ctx.copy_kernel(d_b.ptr, d_a.ptr, nb, 2)
    copy_ms = ctx.copy_kernel(d_b.ptr, d_a.ptr, nb, 10)
    copy_gbs = 2 * nb / (copy_ms * 1e-3) / 1e9
    d_a.free()
    d_b.free()
    # MFMA reference: the dense matrix-core rate per dtype (best of three)
    mfma_peak = {name: round(max(ctx.mfma_peak(d, 4000) for _ in range(3)), 1)
                 for d, name in ((0, "f16"), (1, "bf16"), (2, "f32"))}

    # oracle: argmax match on the first pairs and the CPU baseline, pairs
    # spread over host threads (the oracle is re-entrant C; ctypes releases
    # the GIL), plus one pair on one thread for the latency
    from oracle import oracle as po
    S = min(args.cpu_sample, P)
    match = 0
    cpu_lat_ms = None
    cpu_value = None
    threads = 0
    if S > 0:
        ta = time.perf_counter()
        s0, t0_, _, _ = pairs[0]
        o0 = po.register(t0_, s0, bw=B, n_voxels=N, kdtree=True)
        cpu_lat_ms = (time.perf_counter() - ta) * 1e3
        threads = max(1, min(args.cpu_threads, S))
        with cf.ThreadPoolExecutor(max_workers=threads) as ex:
            tc0 = time.perf_counter()
            outs = list(ex.map(lambda k: po.register(pairs[k][1], pairs[k][0], bw=B, n_voxels=N,
                                                     kdtree=True), range(S)))
            cpu_s = time.perf_counter() - tc0
        cpu_value = S / cpu_s
        for k, o in enumerate(outs):
            match += int(o["rot_index"] == res[k]["rot_index"]
                         and o["trans_index"] == res[k]["trans_index"])
        if other is not None:
            other["match"] = sum(int(o["rot_index"] == other["res"][k]["rot_index"]
                                     and o["trans_index"] == other["res"][k]["trans_index"])
                                 for k, o in enumerate(outs))
        assert o0["rot_index"] == outs[0]["rot_index"]

    variants = variant_throughput(args, pairs, max_pts, tuning) if args.variants else {}
    variants["one_stream_with_grids"] = one_stream
    if no_grids is not None:
        variants[f"{args.streams}_streams_no_grids"] = no_grids

    # argmax match over the committed FP64 oracle fixture (seeds 0..999 of the
    # same workload), both forms; falls back to the timed sample above
    from phaser_amd import parity
    sweep_res = {}
    fx_path = parity.fixture_path(ROOT, args.target_mode)
    other_mode = None
    if args.parity_pairs and os.path.exists(parity.fixture_path(ROOT, 1 - args.target_mode)):
        r = parity.sweep(parity.load(parity.fixture_path(ROOT, 1 - args.target_mode)),
                         precision=args.precision, limit=args.parity_pairs)
        other_mode = {"target_mode": 1 - args.target_mode, "form": FORM_NAMES[args.precision],
                      "matched": r["matched"], "pairs": r["pairs"], "rate": round(r["rate"], 5),
                      "near_ties": r["near_ties"], "mismatches": r["mismatches"][:20]}
        print(f"argmax sweep target_mode {1 - args.target_mode}: {r['matched']}/{r['pairs']}",
              file=sys.stderr, flush=True)
    if args.parity_pairs and os.path.exists(fx_path):
        fx = parity.load(fx_path)
        for prec in ([args.precision, 1 - args.precision] if args.compare_forms
                     else [args.precision]):
            r = parity.sweep(fx, precision=prec, limit=args.parity_pairs)
            sweep_res[prec] = {"matched": r["matched"], "pairs": r["pairs"],
                               "rate": round(r["rate"], 5), "rot_matched": r["rot_matched"],
                               "trans_matched": r["trans_matched"], "near_ties": r["near_ties"],
                               "mismatches": r["mismatches"][:20],
                               "fixture": os.path.relpath(fx_path, ROOT)}
            print(f"argmax sweep {FORM_NAMES[prec]}: {sweep_res[prec]['matched']}/"
                  f"{sweep_res[prec]['pairs']}", file=sys.stderr, flush=True)
    if args.precision in sweep_res:
        argmax_match = sweep_res[args.precision]
    else:
        argmax_match = {"matched": match, "pairs": S, "rate": round(match / max(S, 1), 4),
                        "fixture": None}

    # served forms: a short Poisson sweep of the serving mix per form
    served = None
    if args.serve_rates:
        pool = serve_pool(args)
        rates = [float(x) for x in args.serve_rates.split(",")]
        served = {}
        for prec in ([0, 1] if args.compare_forms else [args.precision]):
            sw, stats = serve_sweep(args, pool, prec, rates, args.serve_duration)
            served[FORM_NAMES[prec]] = dict(serve_summary(sw), server=stats)

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
        "dtype": "fp32" if args.precision == 0 else "fp32 (contractions split fp16)",
        "data": "synthetic (seeded ray-cast scenes, seeds 0..batch-1)",
        "config": {"workload": "B=64 S2 grid 128x128, SO(3) 128^3, voxels 128^3 over +-50 m, "
                               "128x1024-ray scans, target 70% overlap, noise 2 cm, target "
                               + ("re-cast from the moved sensor" if args.target_mode == 1
                                  else "points rigidly moved"),
                   "target_mode": args.target_mode,
                   "pairs_per_step": P, "streams": args.streams, "mean_points_per_cloud": round(P_mean),
                   "form": FORM_NAMES[args.precision], "fft_column_tiles": tuning,
                   "hip_graphs": bool(args.graphs)},
        "p50_ms": round(float(np.percentile(lat, 50)), 4),
        "p99_ms": round(float(np.percentile(lat, 99)), 4),
        "argmax_match": argmax_match,
        "argmax_match_other_target_mode": other_mode,
        "stage_ms_per_step": stage_ms,
        "forms": None if other is None else {
            FORM_NAMES[args.precision]: {"value": round(value, 2),
                                         "ms_per_step": round(elapsed / args.steps * 1e3, 4),
                                         "argmax_match": sweep_res.get(args.precision, {}).get(
                                             "rate", round(match / max(S, 1), 4)),
                                         "stage_ms": {k: stage_ms.get(k) for k in
                                                      ("s2_transform", "so3_synth")}},
            FORM_NAMES[other["precision"]]: {
                "value": round(other["value"], 2), "ms_per_step": round(other["ms_per_step"], 4),
                "argmax_match": sweep_res.get(other["precision"], {}).get(
                    "rate", round(other.get("match", 0) / max(S, 1), 4)),
                "one_stream_vs_this_form_one_stream": round(other["value"] / one_stream["value"], 4),
                "stage_ms": other["stage_ms"]}},
        "variants": variants,
        "served": served,
        "roofline": {"bound": "hbm", "kernel": kname, "stage": dom, "achieved": round(achieved, 2),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 5),
                     "traffic": load_traffic(kname), "bytes_per_launch": int(kbytes),
                     "avg_launch_ms": round(dom_ms, 4),
                     "copy_kernel_gbs": round(copy_gbs, 1),
                     "mfma_measured_tflops": mfma_peak},
        "cpu_baseline": {"value": None if cpu_value is None else round(cpu_value, 4),
                         "unit": "registrations/s", "cores": threads, "kind": "port",
                         "latency_ms_1_thread": None if cpu_lat_ms is None else round(cpu_lat_ms, 1),
                         "host_cpus": os.cpu_count(), "host_cpu_model": host_cpu_model(),
                         "sample": f"{S} pairs of the same workload through the FP64 oracle "
                                   f"(kd-tree NN, own FFT) on {threads} host threads, one pair "
                                   f"per thread at a time"},
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
