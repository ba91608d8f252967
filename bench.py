"""Benchmark: simulated voxels/s on one MI355X, learned + batched against the
plain GPU full scan (the headline metric), or the plain full scan alone.

Default workload (the learned-proposal configuration): a seeded Float32
Gaussian random field 256x256x64 (exponential covariance, ranges 40,40,8),
tile 32x32x8, overlap 1/4, output 384x384x64, tol 0.1, nreal 16, raster path,
2,885,625 candidate patterns.  One plan holds the training image resident in
HBM; all 16 realizations advance in lockstep (batched).  Two legs run on the
same plan with the same uniforms:
  full scan   every tile step scans all candidates exactly (MFMA scan with an
              interval bound, FP64 recheck of the band superset);
  learned     the shortlist of the K best approximate distances (bank scan of
              64-d bf16 encodings) replaces the full scan, reranked in FP64;
              every check_every-th batch (default 16) hands an evenly
              strided sample of check_queries of its queries (default 32, one
              MFMA tile) to the full scan, which decides them and measures
              recall on their exact bands; recall falling below 99 % returns
              the run to the full scan.  With the conv encoder, one more run per encoder checks
              every query of every batch (encoder_ablation_every_batch_checked).
One step = one complete simulation of all realizations (every tile step:
distances, band, weights, draw, min cuts, paste) plus the download of the
output grids to the host.  value = learned voxels/s; the full scan's voxels/s,
the ratio and the recall go in `config.learned_proposal`.  Both legs are timed
the same way (no instrumentation); one further run of each, with HIP events
around the kernels, gives the roofline and kernel shares.

  roofline            k_learn_mscan (the learned leg's bank scan, HBM-bound):
                      algorithmic bytes (bank rows + A2 maps of each mask shape
                      a launch scans) / its HIP-event time; `traffic` from the
                      rocprofv3 FETCH_SIZE / WRITE_SIZE summary in profiles/.
  roofline_full_scan  k_scan_mfma (the full scan): algorithmic FLOPs
                      (2 * candidates * masked voxels per query) / its time,
                      against the FP32 and the bf16 matrix peaks.
  cpu_baseline        the FP64 CPU restatement (oracle/, OpenMP) on the first
                      tile steps of one realization of the same workload.

--workload 2 gives the plain line (3-D Float32 GRF 100x100x50 -> 200x200x50,
tile 20x20x5, full scan only); 1, 3, 4 time the other workloads the same way.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload N]
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
TRAFFIC = {2: os.path.join(ROOT, "profiles", "scan_mfma_grf3d_100x100x50_traffic.md"),
           5: os.path.join(ROOT, "profiles", "bank_scan_grf3d_256x256x64_traffic.md")}

METRIC_LEARNED = "simulated voxels/s on 1 MI355X: learned+batched vs full scan at ≥99% recall"
METRIC_PLAIN = "simulated voxels/s on 1 MI355X: plain GPU full scan"

WORKLOADS = {
    1: "2-D Float64 binary GRF 200x200 (range 20, 30% ones) -> 400x400, tile 30x30, overlap 1/6, tol 0.1, nreal 1",
    2: "plain GPU full scan, 3-D Float32 GRF 100x100x50 (ranges 20,20,5) -> 200x200x50, tile 20x20x5, "
       "overlap 1/5, tol 0.1, nreal 1, raster path",
    3: "batched: 3-facies UInt8 channels 150x150x30, 64 realizations on grids U{100..300}^2 x 30, tile 24x24x6, "
       "overlap 1/4, tol 0.1, one plan, lockstep waves",
    4: "soft + hard: 2-D Float32 GRF 512x512 (range 32), box-blurred soft pair, 500 hard points, -> 1024x1024, "
       "tile 48x48, overlap 1/6, tol 0.05, nreal 32",
    5: "learned + batched vs full scan: 3-D Float32 GRF 256x256x64 (ranges 40,40,8) -> 384x384x64, tile 32x32x8, "
       "overlap 1/4, tol 0.1, nreal 16 in lockstep, 2,885,625 candidates",
}


def measured_traffic(workload):
    """HBM bytes per launch of the roofline kernel from its committed counter
    summary (line 'traffic_bytes_per_launch: <bytes>'), or None."""
    try:
        with open(TRAFFIC[workload]) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch:"):
                    return float(line.split(":", 1)[1])
    except (OSError, KeyError):
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
    """Time the oracle on the first `sample_steps` raster tile steps of one realization."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import iqref

    iqref.build()
    T = w["trainimg"]
    sim = w["simsize"] if "simsize" in w else w["simsizes"][0]
    geo = iqref.geometry(T.shape, w["tilesize"], sim, w["overlap"])
    nsteps = int(np.prod(geo["ntiles"]))
    n = min(sample_steps, nsteps)
    t0 = time.perf_counter()
    iqref.iqsim(T, w["tilesize"], sim, overlap=w["overlap"], tol=w["tol"], nreal=1, seed=1, max_steps=n,
                soft=w.get("soft", ()), hard=w.get("hard"))
    dt = time.perf_counter() - t0
    vox_per_step = float(np.prod(sim)) / nsteps
    return dict(value=vox_per_step * n / dt, unit="voxels/s", cores=iqref.num_threads(), kind="port",
                sample=f"first {n} of {nsteps} tile steps of one realization ({dt:.1f} s; FP64 direct sums, "
                       f"OpenMP over candidates; voxels = output voxels per tile step x steps)")


def timed_leg(plan, U, warmup, steps, buf):
    """Wall time of `steps` complete simulations, each with the download of
    the output grids, after `warmup` untimed ones; also the device time."""
    lib = plan.ctx._L
    for i in range(warmup):
        plan.run(U[i])
        plan.results(grids=True, chosen=False, out=buf)
    lib.iq_ctx_synchronize(plan.ctx.handle)
    dev = 0.0
    t0 = time.perf_counter()
    for i in range(steps):
        plan.run(U[warmup + i])  # returns once the stream has drained
        dev += plan.last_time()
        plan.results(grids=True, chosen=False, out=buf)
    lib.iq_ctx_synchronize(plan.ctx.handle)
    return time.perf_counter() - t0, dev


def profiled_run(plan, u):
    """One more run with HIP events around the scan, bank-scan and cut launches."""
    plan.profile(True)
    plan.run(u)
    dev = plan.last_time()
    scan = plan.scan_stats()
    cut = plan.cut_time()
    bank = plan.bank_stats()
    rounds, levels, ncuts = plan.cut_stats()
    nsel, nsuper, nband = plan.select_stats()
    plan.profile(False)
    return dict(dev=dev, scan=scan, cut=cut, bank=bank, cutwork=(rounds, levels, ncuts), sel=(nsel, nsuper, nband))


def scan_roofline(p, bf16_exact, workload):
    launches, sec, flop, mflop = p["scan"]
    achieved = flop / sec / 1e12 if sec > 0 else 0.0
    peak = PEAK_BF16_TFLOPS if bf16_exact else PEAK_FP32_TFLOPS
    return {
        "bound": "mfma", "achieved": achieved, "peak": peak, "unit": "TFLOP/s", "frac": achieved / peak,
        "traffic": measured_traffic(workload) if workload == 2 else None,
        "kernel": "k_scan_mfma", "launches": launches, "avg_launch_us": sec / max(launches, 1) * 1e6,
        "flop_per_launch": flop / max(launches, 1),
        "frac_of_fp32_peak": achieved / PEAK_FP32_TFLOPS,
        "mfma_bf16_issued_tflops": mflop / sec / 1e12 if sec > 0 else 0.0,
        "mfma_bf16_util": (mflop / sec / 1e12 / PEAK_BF16_TFLOPS) if sec > 0 else 0.0,
    }


def kernel_shares(p):
    cut_n, cut_s = p["cut"]
    rounds, levels, ncuts = p["cutwork"]
    nsel, nsuper, nband = p["sel"]
    dev = max(p["dev"], 1e-12)
    return {
        "device_ms": p["dev"] * 1e3,
        "mincut": {"cuts": ncuts, "push_rounds_per_cut": rounds / max(ncuts, 1),
                   "relabel_passes_per_cut": levels / max(ncuts, 1), "avg_launch_us": cut_s / max(cut_n, 1) * 1e6,
                   "share_of_device_time": cut_s / dev,
                   "bound": "latency (dependent LDS accesses and barriers per push-relabel round)"},
        "scan_share_of_device_time": p["scan"][1] / dev,
        "bank_scan_share_of_device_time": p["bank"][1] / dev,
        "selection": {"queries": nsel, "superset_per_query": nsuper / max(nsel, 1),
                      "band_per_query": nband / max(nsel, 1)},
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--cpu-sample", type=int, default=-1,
                    help="tile steps timed on the CPU oracle (0 = skip; default: 16 at workload 5, the whole "
                         "simulation at workload 2, else skip)")
    ap.add_argument("--copy-mb", type=int, default=0, help="only run the copy used to calibrate HBM counters")
    ap.add_argument("--workload", type=int, default=5, choices=sorted(WORKLOADS),
                    help="5 (default): learned + batched vs the full scan; 1-4: the full scan alone")
    ap.add_argument("--K", type=int, default=2048, help="shortlist size of the learned proposal")
    ap.add_argument("--check-every", type=int, default=16, help="full-scan recall check every N-th batch")
    ap.add_argument("--check-queries", type=int, default=32,
                    help="queries of a check batch (a strided sample) the full scan decides and measures recall on "
                         "(0: all)")
    ap.add_argument("--full-steps", type=int, default=-1, help="timed steps of the full-scan leg (default: --steps)")
    ap.add_argument("--encoder", choices=("conv", "linear"), default="conv",
                    help="learned leg's encoder: the trained convolutional network (default) or the linear one alone")
    ap.add_argument("--conv-steps", type=int, default=200, help="Adam steps of the convolutional encoder")
    ap.add_argument("--conv-lr-head", type=float, default=0.0, help="learning rate of the heads (0: linear heads stay)")
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

    learned = args.workload == 5
    w = synth.workload(args.workload)
    ctx = iqhip.default_context()
    rng = np.random.default_rng(1)
    t0 = time.perf_counter()
    if "simsizes" in w:  # batched realizations, one grid each
        simsizes = w["simsizes"]
        plan, infos = iqhip.make_plan_multi(w["trainimg"], w["tilesize"], simsizes, overlap=w["overlap"],
                                            tol=w["tol"], nreal=1, rng=rng, ctx=ctx)
    else:
        simsizes = [w["simsize"]] * w["nreal"]
        plan, info = iqhip.make_plan(w["trainimg"], w["tilesize"], w["simsize"], overlap=w["overlap"], tol=w["tol"],
                                     nreal=w["nreal"], soft=w.get("soft", ()), hard=w.get("hard"), rng=rng, ctx=ctx)
    plan_s = time.perf_counter() - t0
    full_steps = args.steps if args.full_steps < 0 else args.full_steps
    nU = args.warmup + max(args.steps, full_steps)
    U = [rng.random(sum(plan.real_npath)) for _ in range(nU)]
    # the output grids land in page-locked host memory (a pageable buffer
    # would add a staging copy to every step's download)
    buf = plan.ctx.pinned_empty(sum(int(np.prod(p)) for p in plan.real_padsize), plan.dtype)
    voxels = float(sum(np.prod(z) for z in simsizes))
    t32 = np.asarray(w["trainimg"], dtype=np.float32)
    bf16_exact = bool(((np.ascontiguousarray(t32).view(np.uint32) & 0xFFFF) == 0).all())
    dtype = "f32" if plan.dtype == np.float32 else "f64"

    # full scan
    f_el, f_dev = timed_leg(plan, U, args.warmup, full_steps, buf)
    full_value = voxels * full_steps / f_el
    f_prof = profiled_run(plan, U[0])
    tune_nw, tune_us = plan.scan_tuning()
    res = {
        "metric": METRIC_PLAIN,
        "value": full_value,
        "unit": "voxels/s",
        "n_gpus": 1,
        "steps": full_steps,
        "warmup": args.warmup,
        "ms_per_step": f_el / full_steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": dtype,
        "data": "synthetic",
        "config": {
            "workload": WORKLOADS[args.workload],
            "tile_steps_per_step": int(sum(plan.real_npath)),
            "step": "one complete simulation of every realization + download of the output grids",
            "plan_create_seconds": plan_s,
            "device_ms_per_step": f_dev / full_steps * 1e3,
            "learned_proposal": "off",
            "full_scan_kernels": kernel_shares(f_prof),
            "scan_tuning": {"wavefronts_per_group": {"1_tile": tune_nw[0], "2+_tiles": tune_nw[1]},
                            "us_per_launch": {f"{t}_tile_{wv}_wavefronts": round(v, 1)
                                              for (t, wv), v in sorted(tune_us.items())}},
        },
        "roofline": scan_roofline(f_prof, bf16_exact, args.workload),
    }
    if learned:
        t0 = time.perf_counter()
        plan.learn(nsamples=8192, iters=16, seed=1)
        ctx._L.iq_ctx_synchronize(ctx.handle)
        learn_s = time.perf_counter() - t0
        ablation = None
        conv = None
        if args.encoder == "conv":
            # ablation: the linear encoders alone, one checked run (recall at the same K)
            plan.set_proposal(True, K=args.K, check_every=1, min_recall=0.0)
            plan.run(U[0])
            st_lin = plan.proposal_stats()
            plan.set_proposal(False)
            # training pairs: the encoder runs the full scan itself and takes
            # tile steps' queries as they run
            t0 = time.perf_counter()
            conv = plan.learn_conv(nqueries=640, topn=256, nrandom=64, nhard=64, steps=args.conv_steps, batch=16,
                                   lr=1e-3, lr_head=args.conv_lr_head, seed=1, val_k=args.K)
            conv["seconds"] = time.perf_counter() - t0
            plan.set_proposal(True, K=args.K, check_every=1, min_recall=0.0)
            plan.run(U[0])
            st_conv = plan.proposal_stats()
            ablation = {"K": args.K, "checked_queries": st_conv["queries"],
                        "recall_linear_encoder": st_lin["recall"], "recall_conv_encoder": st_conv["recall"],
                        "fully_recalled_linear": st_lin["fully_recalled"], "fully_recalled_conv": st_conv["fully_recalled"]}
        plan.set_proposal(True, K=args.K, check_every=args.check_every, min_recall=0.99,
                          check_queries=args.check_queries)
        l_el, l_dev = timed_leg(plan, U, args.warmup, args.steps, buf)
        st = plan.proposal_stats()  # of the last timed run (every run draws the same way)
        l_prof = profiled_run(plan, U[0])
        bank_n, bank_s, bank_b = l_prof["bank"]
        achieved = bank_b / bank_s / 1e9 if bank_s > 0 else 0.0
        value = voxels * args.steps / l_el
        res["metric"] = METRIC_LEARNED
        res["value"] = value
        res["steps"] = args.steps
        res["ms_per_step"] = l_el / args.steps * 1e3
        res["roofline_full_scan"] = res.pop("roofline")
        res["roofline"] = {
            "bound": "hbm", "achieved": achieved, "peak": HBM_GBS, "unit": "GB/s", "frac": achieved / HBM_GBS,
            "traffic": measured_traffic(5), "kernel": "k_learn_mscan", "launches": bank_n,
            "avg_launch_us": bank_s / max(bank_n, 1) * 1e6, "bytes_per_launch": bank_b / max(bank_n, 1),
        }
        res["config"]["learned_proposal"] = {
            "K": args.K, "check_every": args.check_every, "check_queries": args.check_queries or "all",
            "min_recall": 0.99,
            "recall": st["recall"], "checked_queries": st["queries"], "fully_recalled_queries": st["fully_recalled"],
            "fell_back": st["fell_back"], "learn_seconds": learn_s,
            "encoder": ("3-layer 3x3x3 conv (1-32-32-1, residual) + 64-d linear heads, trained on the GPU"
                        if args.encoder == "conv" else "64-d linear heads (subspace iteration)"),
            "conv_training": conv, "encoder_ablation_every_batch_checked": ablation,
            "device_ms_per_step": l_dev / args.steps * 1e3,
            "full_scan_value": full_value, "full_scan_ms_per_step": f_el / full_steps * 1e3,
            "speedup_vs_full_scan": value / full_value,
            "kernels": kernel_shares(l_prof),
        }
    plan.close()
    cs = args.cpu_sample
    if cs < 0:
        cs = {5: 16, 2: 2197}.get(args.workload, 0)
    if cs > 0:
        res["cpu_baseline"] = cpu_baseline(w, cs)
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
