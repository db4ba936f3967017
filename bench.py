#!/usr/bin/env python3
"""Throughput of the X-BACH PSF-matching + subtraction path on one MI355X.

Two forms of the same path are measured on seeded synthetic 2048x2048 FP32
pairs that are already resident in HBM (outputs stay in HBM):

* plain: one pair at a time, one step = one full init -> sub pipeline (masks,
  stamp statistics and substamp selection, basis convolution + normal
  equations, the direction test fits, the global fit with its clip-refit loop,
  the spatially varying convolution and the difference image);
* batched: one step = one batch of --batch distinct pairs run by the batch
  runtime (up to --workers pairs in flight, the count per size bucket chosen
  by the batch autotuner before the timed batches, each pair through the same
  bit-identical pipeline); pair 0 of the batch is the plain pair, and its
  outputs are checked to equal the plain run's bit for bit;
* mixed+batched+screened: the same batches with cmv on the matrix cores and
  the learned substamp screen.

Every form does --warmup untimed steps, then times --steps steps.  `value` is
the mixed+batched+screened form's diffs/s when its residual RMS is within 1 %
of the plain form's and of the FP64 oracle's (the metric's comparison); the
other forms' figures are under "forms".  Prints one JSON line.

    python bench.py [--steps K] [--warmup W] [--batch B] [--workers N] [--form both]

cpu_baseline: the CPU restatement of the same path (oracle/, FP64 with the
reference's FP32 steps) on this host's cores: one warm-up, then the median of
--cpu-runs whole diffs of the same pair, with the 95 % rank interval of the
reference's timing rule (tools/exetime.py:6-49).
"""
from __future__ import annotations

import argparse
import json
import os
import pathlib
import sys
import time

# The batch runtime keeps one HIP stream per pair in flight; HIP maps streams
# onto GPU_MAX_HW_QUEUES hardware queues (4 by default), and streams that share
# a queue run one after the other.  Read before the HIP runtime starts.
HW_QUEUES = 16
os.environ["GPU_MAX_HW_QUEUES"] = str(HW_QUEUES)

import numpy as np

ROOT = pathlib.Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "thesis-emx-bach_amd"))
import screen  # noqa: E402
import xbach  # noqa: E402

METRIC = "2048² diffs/sec: plain FP32 vs mixed+batched+screened, at equal residual RMS"
PEAK_FP64_TFLOPS = 78.6   # MI355X dense FP64 (vector and matrix alike)
PEAK_BF16_TFLOPS = 2500.0  # MI355X dense BF16 MFMA
PEAK_HBM_GBS = 8000.0
# HBM bytes per k_conv launch on the 2048x2048 pair: rocprofv3 --pmc FETCH_SIZE
# and WRITE_SIZE of k_conv, each scaled by what the same counter reads for a
# device copy of known size (bench.py --copy-calibration); the summary with
# both commands is profiles/k_conv_2048_traffic.md.
CONV_TRAFFIC_BYTES = 179_734_812
# HBM bytes per global solve (the 23 kernels k_lu_scale, k_lu_begin, 19 x
# k_lu_step, k_lu_backsub), measured the same way; profiles/lu_solve_traffic.md.
LU_TRAFFIC_BYTES = 23_401_549
# HBM bytes per L D L^T solve (one k_ldlt_solve launch, with one refinement
# step), measured the same way; profiles/ldlt_solve_traffic.md.
LDLT_TRAFFIC_BYTES = 2_992_845
LU_N = 292                                   # unknowns of the global system
LU_FLOPS = 2.0 / 3.0 * LU_N ** 3             # ludcmp's multiply-adds, as flops
LDLT_FLOPS = LU_N ** 3 / 3.0 + 8.0 * LU_N ** 2  # L D L^T, substitutions, one refinement

WORKLOADS = {
    "pair2048": ("2048x2048 FP32 pair, 6000 stars (dN/dF ~ F^-1.5, 1e2-1e5), template FWHM 1.8 px, "
                 "science FWHM 2.5->3.5 px ramp in x and y, sky 200 + 0.02x + 0.01y, Poisson noise; "
                 "49-term basis (sigma 0.7/1.5/3.0, degrees 6/4/2), 21x21 kernel, 10x10 stamps, "
                 "31x31 substamps, kernel order 2, background order 1; plain form: FP64 fit and "
                 "FP64 convolution bit-identical to the reference's arithmetic, FP32 images"),
    "pair512": ("512x512 FP32 pair, 300 stars, FWHM 2.0 / 3.0 px, sky 100, Poisson noise; "
                "plain form"),
}


def progress(msg: str) -> None:
    """One line on stderr per finished part of the run (the JSON line stays alone on stdout)."""
    print(f"bench: {msg}", file=sys.stderr, flush=True)


def rank_ci(times):
    """95 % rank interval of the median by the reference's rule
    (tools/exetime.py:6-21: z = (0.50798 + 0.51197) / 2, ranks
    floor((n - z sqrt n) / 2) and ceil(1 + (n + z sqrt n) / 2) of the sorted
    times)."""
    import math

    z = (0.50798 + 0.51197) / 2
    t = sorted(times)
    n = len(t)
    lo = math.floor((n - z * math.sqrt(n)) / 2)
    hi = min(n - 1, math.ceil(1 + (n + z * math.sqrt(n)) / 2))
    return t[lo], t[hi]


STAGES = ["Ini", "SSS", "CMV", "CD", "KSC", "Conv", "Sub", "Fin", "BACH"]


def cpu_baseline(params, tmpl, sci, runs: int, threads: int = 0) -> dict:
    """The oracle on this host: one warm-up, then `runs` whole diffs of the same
    pair; diffs/s of the median run (threads 0: all host cores)."""
    sys.path.insert(0, str(ROOT / "tests"))
    import _oracle  # test infrastructure: the CPU restatement

    threads = threads or int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or os.cpu_count() or 1
    _oracle.run(params, tmpl, sci, threads=threads)  # warm-up
    times, stages = [], []
    for _ in range(runs):
        t0 = time.perf_counter()
        _, _, _, _, ms = _oracle.run(params, tmpl, sci, threads=threads)
        times.append(time.perf_counter() - t0)
        stages.append(ms)
    med = float(np.median(times))
    lo, hi = rank_ci(times)
    st = np.median(np.array(stages), axis=0)
    return {"value": 1.0 / med, "unit": "diffs/s", "cores": threads, "kind": "port",
            "sample": f"median of {runs} whole {tmpl.shape[1]}x{tmpl.shape[0]} diffs of the same "
                      f"pair after one warm-up ({sum(times):.1f} s; OpenMP over stamps and rows, "
                      f"global LU serial)",
            "median_s": med, "ci95_s": [lo, hi],
            "ci_within_10pct": bool(lo >= 0.9 * med and hi <= 1.1 * med),
            "stage_ms": {k: float(v) for k, v in zip(STAGES, st)}}


def fp64_residual(params, tmpl, sci) -> float:
    sys.path.insert(0, str(ROOT / "tests"))
    import _oracle
    import _compare

    _, d, _, _, _ = _oracle.run(params, tmpl, sci, mode=_oracle.MODE_FP64)
    return _compare.residual_rms(d)


def synth_batch(base, n: int) -> list:
    """n distinct pairs of the named workload: seeds base.seed + i (pair 0 is
    the plain pair)."""
    cfgs = []
    for i in range(n):
        c = xbach.SynthCfg.from_buffer_copy(base)
        c.seed = base.seed + i
        cfgs.append(c)
    return xbach.synth_pairs(cfgs, threads=min(16, os.cpu_count() or 1))


def run_plain(ctx, params, w, h, d_t, d_s, steps, warmup):
    for _ in range(warmup):
        ctx.run_device(params, w, h, d_t, d_s)
    ctx.profile(True)
    t0 = time.perf_counter()
    for _ in range(steps):
        ctx.run_device(params, w, h, d_t, d_s)  # returns after the stream drained
    t1 = time.perf_counter()
    prof = ctx.profile_get()
    ctx.profile(False)
    return (t1 - t0) * 1e3 / steps, prof


def upload_pairs(ctx, pairs):
    """Device buffers (tmpl, sci, conv, diff) per pair, inputs uploaded."""
    bufs = []
    for t, s in pairs:
        q = [ctx.buf_alloc(t.nbytes) for _ in range(4)]
        ctx.buf_upload(q[0], t)
        ctx.buf_upload(q[1], s)
        bufs.append(q)
    return bufs


def run_batched(ctx, params, pairs, bufs, workers, steps, warmup, form=0, model=None,
                keep=4, tune=True, flags=0):
    """Time `steps` batches of all pairs (device-resident in and out); returns
    ms per batch, the per-pair results, the diffs of the first `keep` pairs,
    the autotuner's pairs in flight and launch shapes per size bucket (set up before the timed
    batches)."""
    ws = [t.shape[1] for t, _ in pairs]
    hs = [t.shape[0] for t, _ in pairs]
    cols = [[q[k] for q in bufs] for k in range(4)]
    table = None
    shapes = None
    with xbach.Batch(0, workers=workers, flags=flags) as b:
        if model is not None:
            b.screen_load(model)
        b.set_form(form)
        if tune:
            table = b.autotune(params, ws, hs, *cols, reps=2)
            # the tuned launch shapes of each size bucket present
            shapes = {k: b.shapes(k).as_dict() for k in range(4) if table[k]}
        for _ in range(warmup):
            b.run_device(params, ws, hs, *cols)
        t0 = time.perf_counter()
        for _ in range(steps):
            res = b.run_device(params, ws, hs, *cols)
        t1 = time.perf_counter()
    diffs = []
    for i in range(min(keep, len(pairs))):
        d = np.empty_like(pairs[i][0])
        ctx.buf_download(d, bufs[i][3])
        diffs.append(d)
    return (t1 - t0) * 1e3 / steps, res, diffs, table, shapes


def cmv_mfma_rate(ctx, params, w, h, d_t, d_s, reps=10) -> dict:
    """The mixed form's cmv stage (both images' W, Q and B on the matrix cores)
    timed on its own (host clock around the stage, both streams drained): algorithmic flops per the separable
    method (basis 2*(1581+961)*21*49, Q 2*(50*51/2)*961, B 2*50*961 per stamp)
    and the matrix-core flops actually issued: six BF16 part products of the
    512-tap unrolled convolution (W), and the FP64 MFMA products of the 64x64
    normal matrix over 16 chunks of 64 pixels (G')."""
    ctx.set_form(xbach.FORM_MIXED)
    ms = []
    for _ in range(reps):
        ctx.init_device(params, w, h, d_t, d_s)
        ctx.sss()
        ctx.sync()
        t0 = time.perf_counter()
        ctx.cmv()  # the two images' fills run on the context's two streams
        ctx.sync()
        ms.append((time.perf_counter() - t0) * 1e3)
    tr = ctx.trace()
    ctx.set_form(xbach.FORM_PLAIN)
    n = int(tr.stampCount[0]) + int(tr.stampCount[1])
    algo = n * (2.0 * (1581 + 961) * 21 * 49 + 2.0 * (50 * 51 / 2) * 961 + 2.0 * 50 * 961)
    # per stamp: W = 32 pixel tiles x 2 basis tiles x 32 steps of six
    # 32x32x16 BF16 part products; G' = 16 chunks x 16 waves x 4 steps x 4
    # tiles of 16x16x4 FP64 products
    issued_bf16 = n * 6 * (32 * 2 * 32) * (32 * 32 * 16 * 2)
    issued_fp64 = n * (16 * 16 * 4 * 4) * (16 * 16 * 4 * 2)
    t = float(np.median(ms)) * 1e-3
    return {"stamps": n, "ms": t * 1e3, "algorithmic_TFLOPs": algo / t / 1e12,
            "issued_bf16_TFLOPs": issued_bf16 / t / 1e12,
            "issued_fp64_TFLOPs": issued_fp64 / t / 1e12,
            # share of the stage's time the matrix cores would need at peak
            "mfma_busy_at_peak": (issued_bf16 / PEAK_BF16_TFLOPS
                                  + issued_fp64 / PEAK_FP64_TFLOPS) / t / 1e12,
            "peak_bf16_TFLOPs": PEAK_BF16_TFLOPS, "peak_fp64_TFLOPs": PEAK_FP64_TFLOPS}


def screen_crowded(ctx, params, model, n_pairs: int = 4, reps: int = 3) -> dict:
    """The learned screen on its own workload: 3072x3072 crowded fields
    (30000 stars) with 20 % of the selected substamps contaminated (cosmic-ray
    hits, saturation at 6e4, variable sources), seeds 3072001 + i.  Each pair
    runs in the plain form and with the screen (plain arithmetic, so the
    screen's effect is the only difference); time per diff is the median of
    `reps` device-resident runs, passes and counts from the trace."""
    cfgs = [xbach.synth_cfg("crowd3072", seed=3072001 + i) for i in range(n_pairs)]
    ctx.screen_load(model)
    rows = []
    for t, s in screen.contaminated_pairs(ctx, params, cfgs):
        h, w = t.shape
        d_t = ctx.buf_alloc(t.nbytes)
        d_s = ctx.buf_alloc(s.nbytes)
        ctx.buf_upload(d_t, t)
        ctx.buf_upload(d_s, s)
        row = {}
        for name, form in (("plain", xbach.FORM_PLAIN), ("screened", xbach.FORM_SCREEN)):
            ctx.set_form(form)
            ms = []
            for _ in range(reps):
                t0 = time.perf_counter()
                ctx.run_device(params, w, h, d_t, d_s)
                ms.append((time.perf_counter() - t0) * 1e3)
            tr = ctx.trace()
            d = np.empty_like(t)
            ctx.buf_download(d, ctx.device_outputs()[1])
            row[name] = {"ms": float(np.median(ms)), "passes": int(tr.kscIterations),
                         "rms": residual(d), "screened": int(tr.screened),
                         "restored": int(tr.screenRestored), "confirmed": int(tr.screenConfirmed)}
        ctx.set_form(xbach.FORM_PLAIN)
        ctx.buf_free(d_t)
        ctx.buf_free(d_s)
        rows.append(row)
    tot = {k: {f: sum(r[k][f] for r in rows) for f in ("ms", "passes")}
           for k in ("plain", "screened")}
    return {"pairs": len(rows),
            "workload": "3072x3072 crowded field, 30000 stars, 20 % of the selected substamps "
                        "contaminated; plain arithmetic with and without the screen",
            "passes_plain": tot["plain"]["passes"], "passes_screened": tot["screened"]["passes"],
            "ms_per_diff_plain": tot["plain"]["ms"] / len(rows),
            "ms_per_diff_screened": tot["screened"]["ms"] / len(rows),
            "speedup": tot["plain"]["ms"] / tot["screened"]["ms"],
            "screened": sum(r["screened"]["screened"] for r in rows),
            "restored_by_chi2_recheck": sum(r["screened"]["restored"] for r in rows),
            "confirmed_by_chi2_recheck": sum(r["screened"]["confirmed"] for r in rows),
            "residual_rms_rel_max": max(abs(r["screened"]["rms"] - r["plain"]["rms"])
                                        / r["plain"]["rms"] for r in rows)}


def residual(diff) -> float:
    good = diff != np.float32(1e-30)
    return float(np.sqrt(np.mean(diff[good].astype(np.float64) ** 2)))


def same_bits(a, b) -> bool:
    nan = np.isnan(b)
    return bool(np.array_equal(np.isnan(a), nan)
                and np.array_equal(a[~nan].view(np.uint32), b[~nan].view(np.uint32)))


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20,
                    help="timed steps of every form (plain: pipelines; batched: batches)")
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--batch", type=int, default=48, help="pairs per batched step")
    ap.add_argument("--workers", type=int, default=16,
                    help="worker contexts (the most pairs in flight)")
    ap.add_argument("--eager", action="store_true",
                    help="launch the global solves kernel by kernel instead of replaying their "
                         "hipGraphs (for counter collection, which does not follow graph launches)")
    ap.add_argument("--no-tune", action="store_true",
                    help="run all workers' pairs at once instead of the autotuned counts")
    ap.add_argument("--form", default="both", choices=["both", "plain"])
    ap.add_argument("--workload", default="pair2048", choices=sorted(WORKLOADS))
    ap.add_argument("--cpu-runs", type=int, default=15,
                    help="timed whole diffs of the CPU baseline (all cores; half as many, at "
                         "least 10, on one thread)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a device-to-device copy of 256 MiB (counter scale for "
                         "rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE) and exit")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        n = 256 << 20
        src = torch.ones(n // 4, dtype=torch.float32, device="cuda:0")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": n, "copy_bytes_written": n, "launches": 3}))
        return 0
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(the multi-GPU split is a design note, not a built path)", file=sys.stderr)
        return 2

    cfg = xbach.synth_cfg(args.workload)
    tmpl, sci = xbach.synth_pair(cfg)
    h, w = tmpl.shape
    params = xbach.Params.default()

    ctx_flags = xbach.CTX_EAGER if args.eager else 0
    ctx = xbach.Context(0, flags=ctx_flags)
    d_t = ctx.buf_alloc(tmpl.nbytes)
    d_s = ctx.buf_alloc(sci.nbytes)
    ctx.buf_upload(d_t, tmpl)
    ctx.buf_upload(d_s, sci)

    plain_ms, prof = run_plain(ctx, params, w, h, d_t, d_s, args.steps, args.warmup)
    plain_value = 1e3 / plain_ms
    progress(f"plain {plain_value:.1f} diffs/s")
    d_conv, d_diff = ctx.device_outputs()
    conv = np.empty_like(tmpl)
    diff = np.empty_like(tmpl)
    ctx.buf_download(conv, d_conv)
    ctx.buf_download(diff, d_diff)
    rms = residual(diff)

    # the fast form's global solve (L D L^T, one refinement step) timed the
    # same way, one pair at a time with the mixed cmv
    ctx.set_form(xbach.FORM_MIXED | xbach.FORM_LDLT | xbach.FORM_REFINE)
    mixed1_ms, mprof = run_plain(ctx, params, w, h, d_t, d_s, max(5, args.steps // 2),
                                 max(1, args.warmup // 2))
    ctx.set_form(xbach.FORM_PLAIN)
    ldlt_ms, ldlt_n = mprof["solve"]
    ldlt_avg_s = ldlt_ms / max(ldlt_n, 1) * 1e-3
    progress(f"mixed+ldlt single pair {1e3 / mixed1_ms:.1f} diffs/s, solve {ldlt_avg_s * 1e6:.1f} us")

    # roofline of the convolution kernel (the data-parallel hot kernel), from
    # the plain run's HIP events on the stream k_conv runs on: the FP64 441-tap
    # stencil at 882 flop per output pixel is compute-bound (about 70 flop per
    # HBM byte, far above the FP64 ridge of ~10)
    conv_ms, conv_n = prof["conv"]
    conv_avg_s = conv_ms / max(conv_n, 1) * 1e-3
    conv_flops = 2.0 * 441 * (w - 20) * (h - 20)
    conv_tflops = conv_flops / conv_avg_s / 1e12
    kernels = {k: {"ms_per_step": v[0] / args.steps, "launch_groups_per_step": v[1] / args.steps}
               for k, v in prof.items()}
    # the global FP64 solve (ludcmp + lubksb of the 292-unknown system): the
    # largest share of a plain diff.  Each launch group is one whole solve
    # (23 kernels, HIP events on the stream it runs on); its arithmetic is
    # (2/3) 292^3 flops, bound by the latency of 292 dependent pivot columns
    # and of the 42,486-step ordered back substitution, not by FP64 throughput
    solve_ms, solve_n = prof["solve"]
    solve_avg_s = solve_ms / max(solve_n, 1) * 1e-3
    solve_tflops = LU_FLOPS / solve_avg_s / 1e12
    kernels["solve"].update({"ms_per_solve": solve_avg_s * 1e3,
                             "us_per_column": solve_avg_s * 1e6 / LU_N,
                             "achieved_TFLOPs": solve_tflops})
    kernels.pop("sub", None)  # sub is formed in k_conv's epilogue
    kernels["conv"]["achieved_TFLOPs"] = conv_tflops
    # conv + sub bytes (one kernel): algorithmic (FP32 template and science
    # in, FP32 convolved and difference images and the u16 mask out, FP64 tile
    # kernels in) and as measured by the counters
    conv_bytes = 18.0 * w * h + 8.0 * 441 * -(-(w - 20) // 21) * -(-(h - 20) // 21)
    kernels["conv"]["achieved_GBs_algorithmic"] = conv_bytes / conv_avg_s / 1e9
    if (w, h) == (2048, 2048):
        kernels["conv"]["achieved_GBs_measured_traffic"] = CONV_TRAFFIC_BYTES / conv_avg_s / 1e9

    # the same pipeline with host buffers in and out (H2D of both images, D2H of
    # the convolved and difference images, pageable host memory), with the
    # reference's nine-stage -vt breakdown (each stage synchronised)
    for _ in range(2):
        ctx.run(params, tmpl, sci)
    t0 = time.perf_counter()
    nh = max(3, args.steps // 4)
    stage = []
    for _ in range(nh):
        stage.append(ctx.run(params, tmpl, sci)[2])
    host_ms = (time.perf_counter() - t0) * 1e3 / nh
    capture_failures = int(ctx.trace().captureFailures)

    forms = {"plain": {"value": plain_value, "ms_per_diff": plain_ms, "steps": args.steps,
                       "warmup": args.warmup, "residual_rms": rms},
             "mixed_ldlt_single_pair": {"value": 1e3 / mixed1_ms, "ms_per_diff": mixed1_ms,
                                        "solve_us": ldlt_avg_s * 1e6,
                                        "note": "one pair at a time, mixed cmv, L D L^T "
                                                "solves with one refinement step"},
             "plain_host_buffers": {
                 "value": 1e3 / host_ms, "ms_per_diff": host_ms, "steps": nh,
                 "note": "PCIe-inclusive, not the headline value",
                 "stage_ms": {k: float(v) for k, v in zip(STAGES, np.median(stage, axis=0))},
                 "stage_note": "-vt breakdown, a sync after each stage; cd and both "
                               "directions' clip-refit loops run at once, so CD reads 0 and "
                               "KSC holds both"}}
    value, ms_per_step, steps, warmup, form = plain_value, plain_ms, args.steps, args.warmup, "plain"
    ref_rms = None
    if not args.no_cpu:
        ref_rms = fp64_residual(params, tmpl, sci)
        progress("FP64 oracle residual done")
    if args.form == "both":
        pairs = synth_batch(cfg, args.batch)
        assert np.array_equal(pairs[0][0], tmpl) and np.array_equal(pairs[0][1], sci)
        bufs = upload_pairs(ctx, pairs)
        keep = 8
        b_ms, res, dplain, btab, bshp = run_batched(ctx, params, pairs, bufs, args.workers,
                                              args.steps, args.warmup, keep=keep,
                                              tune=not args.no_tune, flags=ctx_flags)
        ok = all(r.status == 0 for r in res)
        b_value = args.batch * 1e3 / b_ms
        progress(f"batched {b_value:.1f} diffs/s")
        forms["batched"] = {
            "value": b_value, "ms_per_batch": b_ms, "pairs_per_batch": args.batch,
            "workers": args.workers, "pairs_in_flight_per_bucket": btab,
            "launch_shapes_per_bucket": bshp,
            "hw_queues": HW_QUEUES,
            "steps": args.steps, "warmup": args.warmup, "all_pairs_ok": ok,
            "pair0_bit_identical_to_plain": same_bits(dplain[0], diff),
            "pair_ms_mean": float(np.mean([r.ms for r in res])),
            "capture_failures": int(sum(r.captureFailures for r in res)),
            "speedup_vs_plain": b_value / plain_value}
        # the same batches in lockstep groups: every global solve of a group
        # (all its pairs' direction trials, then each clip-refit pass) is one
        # batched solve
        l_ms, lres, dls, ltab, lshp = run_batched(ctx, params, pairs, bufs, args.workers,
                                            args.steps, args.warmup, keep=keep,
                                            tune=not args.no_tune,
                                            flags=ctx_flags | xbach.BATCH_LOCKSTEP)
        l_value = args.batch * 1e3 / l_ms
        progress(f"lockstep {l_value:.1f} diffs/s")
        forms["batched_lockstep"] = {
            "value": l_value, "ms_per_batch": l_ms, "pairs_per_batch": args.batch,
            "pairs_per_group_per_bucket": ltab, "steps": args.steps, "warmup": args.warmup,
            "all_pairs_ok": all(r.status == 0 for r in lres),
            "bit_identical_to_batched": all(same_bits(a, b) for a, b in zip(dls, dplain)),
            "speedup_vs_batched": l_value / b_value}
        # mixed precision + batched + learned screen: cmv on the matrix cores,
        # the global solves by L D L^T with one FP64 refinement step, the screen
        model = np.load(screen.MODEL_PATH) if screen.MODEL_PATH.exists() else None
        scr = xbach.FORM_SCREEN if model is not None else 0
        fform = xbach.FORM_MIXED | xbach.FORM_LDLT | xbach.FORM_REFINE | scr
        f_ms, fres, dfull, ftab, fshp = run_batched(ctx, params, pairs, bufs, args.workers,
                                              args.steps, args.warmup, form=fform, model=model,
                                              keep=keep, tune=not args.no_tune,
                                              flags=ctx_flags)
        # the same with the global solves in Numerical Recipes' order (ludcmp +
        # lubksb, the plain form's solve)
        n_ms, nres, dnr, ntab, nshp = run_batched(ctx, params, pairs, bufs, args.workers,
                                            args.steps, args.warmup,
                                            form=xbach.FORM_MIXED | scr, model=model, keep=keep,
                                            tune=not args.no_tune, flags=ctx_flags)
        n_value = args.batch * 1e3 / n_ms
        progress(f"mixed+batched+screened, ludcmp solves {n_value:.1f} diffs/s")
        forms["mixed_batched_screened_ludcmp"] = {
            "value": n_value, "ms_per_batch": n_ms, "pairs_per_batch": args.batch,
            "pairs_in_flight_per_bucket": ntab, "steps": args.steps, "warmup": args.warmup,
            "all_pairs_ok": all(r.status == 0 for r in nres),
            "speedup_vs_batched_plain": n_value / b_value}
        for q in bufs:
            for d in q:
                ctx.buf_free(d)
        del pairs
        fok = all(r.status == 0 for r in fres)
        f_value = args.batch * 1e3 / f_ms
        progress(f"mixed+batched+screened {f_value:.1f} diffs/s")
        rel = [abs(residual(a) - residual(b)) / residual(b) for a, b in zip(dfull, dplain)]
        full = {
            "value": f_value, "ms_per_batch": f_ms, "pairs_per_batch": args.batch,
            "workers": args.workers, "pairs_in_flight_per_bucket": ftab,
            "launch_shapes_per_bucket": fshp,
            "form": "mixed+ldlt+refine" + ("+screened" if model is not None else ""),
            "steps": args.steps, "warmup": args.warmup, "all_pairs_ok": fok,
            "residual_rms_rel_vs_plain_max": float(max(rel)), "pairs_checked": len(rel),
            "clip_refit_passes_mean": float(np.mean([r.kscIterations for r in fres])),
            "clip_refit_passes_mean_plain": float(np.mean([r.kscIterations for r in res])),
            "speedup_vs_plain": f_value / plain_value}
        if ref_rms is not None:
            full["residual_rms_rel_vs_fp64_oracle_pair0"] = abs(residual(dfull[0]) - ref_rms) / ref_rms
        equal_rms = fok and max(rel) <= 0.01 and (
            ref_rms is None or full["residual_rms_rel_vs_fp64_oracle_pair0"] <= 0.01)
        full["equal_residual_rms"] = bool(equal_rms)
        forms["mixed_batched_screened"] = full
        full["speedup_vs_batched_plain"] = f_value / b_value
        if equal_rms:
            value, ms_per_step, steps, warmup, form = (f_value, f_ms, args.steps, args.warmup,
                                                       full["form"] + "+batched")
        elif ok and forms["batched"]["pair0_bit_identical_to_plain"]:
            value, ms_per_step, steps, warmup, form = (b_value, b_ms, args.steps, args.warmup,
                                                       "batched")
        forms["cmv_mixed_mfma"] = cmv_mfma_rate(ctx, params, w, h, d_t, d_s)
        if model is not None:
            forms["screen_crowded_3072"] = screen_crowded(ctx, params, model)

    line = {
        "metric": METRIC,
        "value": value,
        "unit": "diffs/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": WORKLOADS[args.workload], "name": args.workload,
                   "form": form, "w": w, "h": h, "seed": int(cfg.seed),
                   "pairs_per_step": 1 if form == "plain" else args.batch},
        # the dominant kernel group: the global FP64 solve, on the FP64 peak
        # (78.6 TFLOP/s, the same for the vector units it runs on and the
        # matrix cores); it is latency-bound, so its fraction is tiny
        # "mfma" is the compute-bound class of the line's schema; the solve's
        # arithmetic is FP64 VALU, whose MI355X peak equals the FP64 MFMA peak
        "roofline": {"bound": "mfma", "compute_unit": "fp64 valu",
                     "kernel": "global FP64 LU solve (k_lu_scale, k_lu_begin, "
                     "19 x k_lu_step, k_lu_backsub; FP64 VALU)",
                     "achieved": solve_tflops, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                     "frac": solve_tflops / PEAK_FP64_TFLOPS,
                     "traffic": LU_TRAFFIC_BYTES,
                     "algorithmic": f"(2/3)*292^3 = {LU_FLOPS:.4g} flop per solve",
                     "us_per_solve": solve_avg_s * 1e6, "us_per_column": solve_avg_s * 1e6 / LU_N,
                     "limiter": "latency: 292 dependent pivot columns and a 42,486-step "
                                "ordered back substitution (bit-identical to ludcmp/lubksb)"},
        # the fast forms' global solve: L D L^T (n^3 / 3 flop, the trailing
        # updates on the FP64 matrix cores) plus the substitutions and one
        # refinement step (8 n^2 flop), one launch per system
        "roofline_ldlt": {"bound": "mfma", "compute_unit": "fp64 mfma",
                          "kernel": "k_ldlt_solve (L D L^T + substitution + one refinement "
                                    "step, one workgroup)",
                          "achieved": LDLT_FLOPS / ldlt_avg_s / 1e12, "peak": PEAK_FP64_TFLOPS,
                          "unit": "TFLOP/s",
                          "frac": LDLT_FLOPS / ldlt_avg_s / 1e12 / PEAK_FP64_TFLOPS,
                          "traffic": LDLT_TRAFFIC_BYTES,
                          "algorithmic": f"292^3/3 + 8*292^2 = {LDLT_FLOPS:.4g} flop per solve",
                          "us_per_solve": ldlt_avg_s * 1e6,
                          "us_per_solve_ludcmp": solve_avg_s * 1e6},
        # the data-parallel stencil: FP64 VALU multiply-adds (unfused, in the
        # reference's term order)
        "roofline_conv": {"bound": "fp64 valu", "kernel": "k_conv",
                          "achieved": conv_tflops, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                          "frac": conv_tflops / PEAK_FP64_TFLOPS,
                          "traffic": CONV_TRAFFIC_BYTES if (w, h) == (2048, 2048) else None,
                          "algorithmic": f"2*441*(W-20)*(H-20) = {conv_flops:.4g} flop per "
                                         "launch"},
        "capture_failures": capture_failures,
        "forms": forms,
        "kernels": kernels,
        "residual_rms": rms,
    }
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(params, tmpl, sci, args.cpu_runs)
        progress(f"cpu baseline {line['cpu_baseline']['value']:.3f} diffs/s")
        # the same restatement on one host thread
        line["cpu_baseline_1thread"] = cpu_baseline(params, tmpl, sci,
                                                    max(10, args.cpu_runs // 2), 1)
        line["residual_rms_fp64_oracle"] = ref_rms
        line["residual_rms_rel_diff"] = abs(rms - ref_rms) / ref_rms
        line["gpu_vs_cpu"] = value / line["cpu_baseline"]["value"]
        line["plain_gpu_vs_cpu"] = plain_value / line["cpu_baseline"]["value"]
    ctx.buf_free(d_t)
    ctx.buf_free(d_s)
    ctx.close()
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())
