"""Benchmark: simulated voxels/s of the plain GPU full scan on one MI355X.

Workload (the plain-GPU configuration of the metric): a seeded Float32
Gaussian random field 100x100x50 (exponential covariance, ranges 20,20,5),
tile 20x20x5, overlap 1/5 per axis, output 200x200x50, tol 0.1, nreal 1,
raster path, no soft or hard data.  One step = one complete simulation (all
2,197 tile steps: distance scan, band pick, tau weights, draw, min cuts,
paste) with the training image already resident in HBM.

Also reported:
  roofline      the MFMA distance scan k_scan_mfma (dominant kernel), timed
                live with HIP events on the library's stream around every scan
                launch during the timed steps; algorithmic FLOPs = 2 *
                candidates * masked voxels per query, summed over the queries
                of a launch.  `traffic` (HBM bytes per launch) comes from the
                rocprofv3 FETCH_SIZE / WRITE_SIZE summary in profiles/, scaled
                by the same counters on a copy of known size (--copy-mb).
  cpu_baseline  the FP64 CPU restatement (oracle/, OpenMP) on a bounded sample
                of the same workload: the first tile steps of one realization.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--cpu-sample N]
       python bench.py --copy-mb M    (device copy of M MiB only: counter calibration)
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "imagequilting.jl_amd"))

PEAK_FP32_TFLOPS = 157.3  # MI355X dense FP32 matrix peak (published)
PEAK_BF16_TFLOPS = 2500.0  # MI355X dense bf16 MFMA peak (published)
HBM_GBS = 8000.0
TRAFFIC_SUMMARY = os.path.join(ROOT, "profiles", "scan_mfma_grf3d_100x100x50_traffic.md")


WORKLOADS = {
    1: "2-D Float64 binary GRF 200x200 (range 20, 30% ones) -> 400x400, tile 30x30, overlap 1/6, tol 0.1, nreal 1",
    2: "plain GPU full scan, 3-D Float32 GRF 100x100x50 (ranges 20,20,5) -> 200x200x50, tile 20x20x5, "
       "overlap 1/5, tol 0.1, nreal 1, raster path",
    3: "batched: 3-facies UInt8 channels 150x150x30, 64 realizations on grids U{100..300}^2 x 30, tile 24x24x6, "
       "overlap 1/4, tol 0.1, one plan, lockstep waves",
    4: "soft + hard: 2-D Float32 GRF 512x512 (range 32), box-blurred soft pair, 500 hard points, -> 1024x1024, "
       "tile 48x48, overlap 1/6, tol 0.05, nreal 32",
    5: "3-D Float32 GRF 256x256x64 (ranges 40,40,8) -> 384x384x64, tile 32x32x8, overlap 1/4, tol 0.1, nreal 16, "
       "full scan (no learned proposal)",
}


def measured_traffic():
    """HBM bytes per k_scan_mfma launch from the committed counter summary
    (line 'traffic_bytes_per_launch: <bytes>'), or None."""
    try:
        with open(TRAFFIC_SUMMARY) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch:"):
                    return float(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def copy_calibration(mib: int):
    """Device-to-device copy of `mib` MiB, five times: the reference kernel
    for scaling FETCH_SIZE / WRITE_SIZE (each copy reads and writes mib MiB)."""
    import torch
    n = mib << 20
    a = torch.empty(n, dtype=torch.uint8, device="cuda:0").fill_(1)
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    for _ in range(5):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_each_way": n, "copies": 5}), flush=True)


def cpu_baseline(w, sample_steps: int):
    """Time the oracle on the first `sample_steps` tile steps (one realization)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import iqref

    iqref.build()
    T = w["trainimg"]
    geo = iqref.geometry(T.shape, w["tilesize"], w["simsize"], w["overlap"])
    nsteps = int(np.prod(geo["ntiles"]))
    # a sub-grid covering the first `sample_steps` raster tiles along x/y rows
    t0 = time.perf_counter()
    iqref.iqsim(T, w["tilesize"], w["simsize"], overlap=w["overlap"], tol=w["tol"], nreal=1, seed=1,
                max_steps=sample_steps)
    dt = time.perf_counter() - t0
    vox_per_step = float(np.prod(w["simsize"])) / nsteps
    return dict(value=vox_per_step * min(sample_steps, nsteps) / dt, unit="voxels/s", cores=iqref.num_threads(),
                kind="port",
                sample=f"first {min(sample_steps, nsteps)} of {nsteps} raster tile steps of one realization "
                       f"({dt:.1f} s, FP64 direct sums, OpenMP over candidates)")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--cpu-sample", type=int, default=2197,
                    help="tile steps timed on the CPU (0 = skip); 2197 = the whole simulation")
    ap.add_argument("--copy-mb", type=int, default=0, help="only run the copy used to calibrate HBM counters")
    ap.add_argument("--workload", type=int, default=2, choices=sorted(WORKLOADS),
                    help="2 (default): the plain full-scan workload; 1, 3, 4, 5: the other workloads")
    ap.add_argument("--learned", action="store_true",
                    help="also time the learned proposal (shortlist + exact rerank) on the same plan and report "
                         "it as the value, with the full scan beside it")
    ap.add_argument("--K", type=int, default=4096, help="shortlist size of the learned proposal")
    ap.add_argument("--check-every", type=int, default=16, help="full-scan recall check every N-th batch")
    args = ap.parse_args()
    if args.copy_mb > 0:
        copy_calibration(args.copy_mb)
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: runs on one GPU only; the multi-GPU layout is a design, not implemented "
              "(realizations shard across GPUs with no communication).", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import iqhip
    from iqhip import synth

    w = synth.workload(args.workload)
    ctx = iqhip.default_context()
    rng = np.random.default_rng(1)
    if "simsizes" in w:  # batched realizations, one grid each
        simsizes = w["simsizes"]
        plan, infos = iqhip.make_plan_multi(w["trainimg"], w["tilesize"], simsizes, overlap=w["overlap"],
                                            tol=w["tol"], nreal=1, rng=rng, ctx=ctx)
    else:
        simsizes = [w["simsize"]] * w["nreal"]
        plan, info = iqhip.make_plan(w["trainimg"], w["tilesize"], w["simsize"], overlap=w["overlap"], tol=w["tol"],
                                     nreal=w["nreal"], soft=w.get("soft", ()), hard=w.get("hard"), rng=rng, ctx=ctx)
        infos = [info]
    U = [rng.random(sum(plan.real_npath)) for _ in range(args.warmup + args.steps)]
    for i in range(args.warmup):
        plan.run(U[i])
    plan.profile(True)
    lib = ctx._L
    lib.iq_ctx_synchronize(ctx.handle)
    t0 = time.perf_counter()
    dev_time = 0.0
    launches, scan_sec, scan_flop, mfma_flop = 0, 0.0, 0.0, 0.0
    cut_launches, cut_sec = 0, 0.0
    for i in range(args.steps):
        plan.run(U[args.warmup + i])  # returns after the stream has drained
        dev_time += plan.last_time()
        n, s, f, mf = plan.scan_stats()
        launches, scan_sec, scan_flop, mfma_flop = launches + n, scan_sec + s, scan_flop + f, mfma_flop + mf
        cn, cs = plan.cut_time()
        cut_launches, cut_sec = cut_launches + cn, cut_sec + cs
    lib.iq_ctx_synchronize(ctx.handle)
    elapsed = time.perf_counter() - t0
    cut_rounds, cut_levels, ncuts = plan.cut_stats()
    nsel, nsuper, nband = plan.select_stats()
    tune_nw, tune_us = plan.scan_tuning()
    voxels = float(sum(np.prod(z) for z in simsizes)) * args.steps
    achieved = scan_flop / scan_sec / 1e12 if scan_sec > 0 else 0.0
    # training images whose values are exact in bf16 (facies codes, binary)
    # are scanned with one bf16 product: price them against the bf16 peak
    t32 = np.asarray(w["trainimg"], dtype=np.float32)
    bf16_exact = bool(((np.ascontiguousarray(t32).view(np.uint32) & 0xFFFF) == 0).all())
    peak = PEAK_BF16_TFLOPS if bf16_exact else PEAK_FP32_TFLOPS
    res = {
        "metric": "simulated voxels/s on 1 MI355X: learned+batched vs full scan at ≥99% recall",
        "value": voxels / elapsed,
        "unit": "voxels/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if plan.dtype == np.float32 else "f64",
        "data": "synthetic",
        "config": {
            "workload": WORKLOADS[args.workload],
            "tile_steps_per_step": int(sum(plan.real_npath)),
            "device_ms_per_step": dev_time / args.steps * 1e3,
            "learned_proposal": "off",
            "mincut": {"cuts": ncuts, "push_rounds_per_cut": cut_rounds / max(ncuts, 1),
                       "relabel_passes_per_cut": cut_levels / max(ncuts, 1),
                       "avg_launch_us": cut_sec / max(cut_launches, 1) * 1e6,
                       "share_of_device_time": cut_sec / max(dev_time, 1e-12),
                       "bound": "latency (dependent LDS accesses and barriers per push-relabel round)"},
            "scan_share_of_device_time": scan_sec / max(dev_time, 1e-12),
            "selection": {"queries": nsel, "superset_per_query": nsuper / max(nsel, 1),
                          "band_per_query": nband / max(nsel, 1)},
            "scan_tuning": {"wavefronts_per_group": {"1_tile": tune_nw[0], "2+_tiles": tune_nw[1]},
                            "us_per_launch": {f"{t}_tile_{wv}_wavefronts": round(v, 1)
                                              for (t, wv), v in sorted(tune_us.items())}},
        },
        "roofline": {
            "bound": "mfma",
            "achieved": achieved,
            "peak": peak,
            "unit": "TFLOP/s",
            "frac": achieved / peak,
            "traffic": measured_traffic() if args.workload == 2 else None,
            "kernel": "k_scan_mfma",
            "launches": launches,
            "avg_launch_us": scan_sec / max(launches, 1) * 1e6,
            "flop_per_launch": scan_flop / max(launches, 1),
            "mfma_bf16_issued_tflops": mfma_flop / scan_sec / 1e12 if scan_sec > 0 else 0.0,
            "mfma_bf16_util": (mfma_flop / scan_sec / 1e12 / PEAK_BF16_TFLOPS) if scan_sec > 0 else 0.0,
        },
    }
    if args.learned:
        # learned + batched: same plan, same uniforms; the full scan above is the baseline
        t0 = time.perf_counter()
        plan.learn(nsamples=8192, iters=16, seed=1)
        lib.iq_ctx_synchronize(ctx.handle)
        learn_s = time.perf_counter() - t0
        plan.set_proposal(True, K=args.K, check_every=args.check_every, min_recall=0.99)
        plan.profile(False)
        for i in range(args.warmup):
            plan.run(U[i])
        t0 = time.perf_counter()
        members = hits = queries = full = 0
        fell = False
        ldev = 0.0
        for i in range(args.steps):
            plan.run(U[args.warmup + i])
            ldev += plan.last_time()
            st = plan.proposal_stats()
            members += st["band_members"]; hits += st["in_shortlist"]
            queries += st["queries"]; full += st["fully_recalled"]; fell |= st["fell_back"]
        lib.iq_ctx_synchronize(ctx.handle)
        lel = time.perf_counter() - t0
        full_scan_value = res["value"]
        res["value"] = voxels / lel
        res["ms_per_step"] = lel / args.steps * 1e3
        res["config"]["learned_proposal"] = {
            "K": args.K, "check_every": args.check_every, "min_recall": 0.99,
            "recall": hits / members if members else None, "checked_queries": queries,
            "fully_recalled_queries": full, "fell_back": fell, "learn_seconds": learn_s,
            "device_ms_per_step": ldev / args.steps * 1e3,
            "full_scan_value": full_scan_value, "speedup_vs_full_scan": res["value"] / full_scan_value,
        }
    plan.close()
    if args.cpu_sample > 0 and args.workload == 2:
        res["cpu_baseline"] = cpu_baseline(w, args.cpu_sample)
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
