"""Benchmark: pair-site posteriors/s of the blockwise decoder on one MI355X.

Workload (default ``chromosome``): 356 haplotypes (all 63,190 pairs) x
402,853 segregating sites, positions uniform on [0, 51 Mb), 1/k spectrum,
seed 7; ``infer_blockwise`` mean_only semantics with 2048-site flanks and
the core block size the auto-sizer picks from free device memory. One step
is one decode of every pair over every block, with the packed cohort, pair
ids and output resident in HBM when the timed region starts.

Prints one JSON line (rank 0). Besides the throughput it reports:
* ``roofline``: algorithmic HBM bytes of the decode kernel per launch over
  the kernel's event-timed duration, against the 8 TB/s HBM peak;
* ``cpu_baseline``: the CPU oracle (FP32 restatement of the same recursion,
  OpenMP over pairs) timed on this host on a bounded pair sample over all
  sites (``--cpu-pairs``, 8,000 by default), reported per pair-site;
* ``checks``: the sampled pairs decoded alone on the GPU are bit-identical
  to the oracle, and to the same pairs' columns of the full run.
"""

from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0

# HBM bytes per decode launch pair measured with rocprofv3 FETCH_SIZE and
# WRITE_SIZE passes (scaled by the calibration copy kernel: x2.000 read,
# x1.000 write), for one workload and launch shape; the summary with the
# commands is profiles/fb_decode_chromosome_counters.md. Reported as
# roofline.traffic only when a run matches that workload and shape.
MEASURED_TRAFFIC = {
    ("chromosome", 8, 1, "static", False): 171.36e9,
    ("chromosome", 8, 1, "persistent", False): 169.79e9,
}

WORKLOADS = {
    "chromosome": dict(n=356, S=402_853, span=51_000_000, seed=7, flank=2048, fp64_pairs=256,
                       label="356 haplotypes (63,190 pairs) x 402,853 sites, 1/k spectrum, "
                             "infer_blockwise mean_only, auto core + 2048-site flanks"),
    "biobank": dict(n=1000, S=50_000, span=7_500_000, seed=17, flank=2048, fp64_pairs=1000, host_stream=True,
                    label="1,000 haplotypes (499,500 pairs) x 50,000 sites, 1/k spectrum, "
                          "infer_blockwise mean_only, auto core + 2048-site flanks"),
    "ragged": dict(n=2000, S=262_144, span=262_144 * 150, seed=11, flank=2048, core=32_768, n_jobs=64,
                   label="64 ragged jobs (sites logU[2k,200k] x pairs logU[100,20k]) over a 2,000-haplotype "
                         "1/k cohort at 1 SNP/150 bp, seed 11; 32,768-site cores + 2048-site flanks, mean_only"),
    "serving": dict(n=6404, S=1_000_000, span=150_000_000, seed=13, arena_gb=96.0, ck_gb=8.0, core=8192,
                    rates=(50.0, 200.0, 800.0, 3200.0), seconds=4.0, max_queries=600,
                    label="resident 6,404 x 1,000,000 bitpacked 1/k cohort (seed 13, 150 bp/site); Poisson "
                          "queries: 4-200 haplotypes (all pairs) x 50 kb-5 Mb regions (log-uniform), "
                          "8,192-site cores + 2048-site flanks (run_fb_blockwise defaults), mean_only"),
}


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def algorithmic_bytes(P, S, n, K):
    """Per decode: 4 B mean per pair-site, 16/K B of checkpoint write+read per
    pair-site, packed rows, positions-derived descriptors and pair ids once."""
    W = (S + 63) // 64
    return P * S * (4.0 + 16.0 / K) + n * W * 8 + 8 * S + 8 * P


def sample_checks(args, ctx, G, pos, pairs, P, S, core, flank, Ne, kmu, krho, d_out, stream):
    """Decode a bounded random pair sample alone on the GPU; compare it with
    the same columns of the full run and with the CPU oracle (timed: the
    cpu_baseline)."""
    import ctypes

    from gamma_smc_cu_amd._core import check, lib
    from gamma_smc_cu_amd.device import DeviceArray, run_fb_device
    from gamma_smc_cu_amd.flow_field import DEFAULT_PATH

    rng = np.random.default_rng(123)
    m = min(args.cpu_pairs, P)
    sample = np.sort(rng.choice(P, size=m, replace=False))
    sp = [tuple(map(int, pairs[k])) for k in sample]
    checks = {}
    d_spi = DeviceArray.from_host(pairs[sample, 0].copy())
    d_spj = DeviceArray.from_host(pairs[sample, 1].copy())
    d_sout = DeviceArray((S, m), np.float32)
    run_fb_device(ctx, d_spi, d_spj, m, core, flank, 0, d_sout, stream=stream)
    stream.synchronize()
    alone = d_sout.to_host()
    rows = sorted({0, core - 1, min(core, S - 1), S // 2, S - 1})
    full_rows = np.empty((len(rows), P), np.float32)
    for r_i, r in enumerate(rows):
        tmp = np.empty(P, np.float32)
        check(lib().gsmc_memcpy_dtoh(ctypes.c_void_p(tmp.ctypes.data), d_out.offset(r * P), P * 4))
        full_rows[r_i] = tmp
    checks["alone_vs_mix_bit_exact"] = bool(np.array_equal(full_rows[:, sample].view(np.uint32),
                                                           alone[rows].view(np.uint32)))
    checks["finite_positive"] = bool(np.isfinite(full_rows).all() and (full_rows > 0).all())

    cpu = None
    if not args.no_cpu_baseline:
        from oracle import oracle as O

        T = O.tables_f32(DEFAULT_PATH, 1000, Ne, kmu, krho)
        cores = int(os.environ.get("OMP_NUM_THREADS", "0")) or len(os.sched_getaffinity(0))
        t1 = time.perf_counter()
        ref = O.fb_blockwise_f32(G, pos, sp, T, Ne, core, flank)
        cpu_s = time.perf_counter() - t1
        checks["oracle_bit_exact_sample"] = bool(np.array_equal(ref["mean"].view(np.uint32), alone.view(np.uint32)))
        cpu = {"value": m * S / cpu_s, "unit": "pair-sites/s", "cores": cores, "kind": "port",
               "sample": f"{m} random pairs x all {S} sites (same blocks), {cpu_s:.1f}s; "
                         f"full job extrapolates to {cpu_s * P / m:.0f}s"}
        log(f"cpu baseline: {cpu['value']:.3e} pair-sites/s on {cores} threads ({cpu_s:.1f}s)")

    fp64 = None
    k = min(args.fp64_pairs, m)
    if k > 0:
        from oracle import oracle as O

        T64 = O.tables_f64(DEFAULT_PATH, 1000, Ne, kmu, krho)
        t1 = time.perf_counter()
        ref64 = O.fb_blockwise_f64(G, pos, sp[:k], T64, Ne, core, flank)
        fp64 = O.log_error_report(alone[:, :k], ref64["mean"])
        fp64["sample"] = f"{k} random pairs x all {S} sites"
        fp64["tolerance"] = "max |log mean32 - log mean64| <= 1e-2, median <= 1e-3"
        checks["fp64_error_within_tolerance"] = bool(fp64["max"] <= 1e-2 and fp64["median"] <= 1e-3)
        log(f"fp64 error: max {fp64['max']:.2e} median {fp64['median']:.2e} ({time.perf_counter() - t1:.1f}s)")

    return checks, cpu, fp64, (sample, alone)


def host_stream_run(ctx, pairs, core, flank, sample, alone):
    """The PCIe-inclusive path infer_blockwise uses: decode into host arrays
    through two device output slots and pinned staging buffers (device
    decode, D2H copies and the host scatter overlap). Returns its rate and
    checks the sampled columns against the device-resident run."""
    P, S = pairs.shape[0], alone.shape[0]
    t0 = time.perf_counter()
    r = ctx.run_fb_blockwise(pairs, core_block_sites=core, flank_sites=flank, pair_batch_size=-1, max_streams=2)
    dt = time.perf_counter() - t0
    same = bool(np.array_equal(r["mean"][:, sample].view(np.uint32), alone.view(np.uint32)))
    gb = 4.0 * P * S / 1e9
    log(f"host stream: {dt:.2f}s for {gb:.1f} GB of output ({gb / dt:.1f} GB/s)")
    return {"seconds": dt, "pair_sites_per_s": float(P) * S / dt, "output_gb": gb, "output_gb_per_s": gb / dt,
            "sample_matches_device_run": same}


def run_ragged(args, wl):
    """Batched ragged form: every job's (pair tile x block window) items in
    one launch, against the plain mode (one launch per job, each waited
    for). Both write the same buffers, and the outputs must be identical."""
    import ctypes

    from gamma_smc_cu_amd import _core, autotune, synthetic
    from gamma_smc_cu_amd.batch import Job, Staged, Workspace
    from gamma_smc_cu_amd.device import Event, Stream, synchronize
    from gamma_smc_cu_amd.flow_field import DEFAULT_PATH
    from gamma_smc_cu_amd.infer import _estimate_scaled_params

    n, S = wl["n"], wl["S"]
    t0 = time.perf_counter()
    G, pos = synthetic.cohort(n, S, wl["span"], wl["seed"])
    raw = synthetic.ragged_jobs(n, S, wl["n_jobs"], wl["seed"])
    log(f"cohort {n} x {S} and {len(raw)} jobs in {time.perf_counter() - t0:.1f}s")
    Ne = 1e4
    kmu, krho = _estimate_scaled_params(G, pos, 1.25e-8, 1e-8, Ne)
    _core.set_device(0)
    ctx = _core.FlowContext(G, pos, Ne, kmu, krho, DEFAULT_PATH, 0)
    jobs = [Job(lo, hi, p, core_block_sites=wl["core"], flank_sites=wl["flank"]) for lo, hi, p in raw]
    pair_sites = float(sum(j.n_sites * len(j.pairs) for j in jobs))
    st = Staged(jobs, mean_only=True)
    ws = Workspace(ctx)
    stream = Stream()
    all_arr, n_all = st.structs()
    one = [st.structs([k]) for k in range(len(jobs))]

    def batched():
        e0, e1 = Event(), Event()
        e0.record(stream)
        ws.launch(all_arr, n_all, st.flags, stream)
        e1.record(stream)
        stream.synchronize()
        return e0.elapsed_ms(e1)

    def plain():
        e0, e1 = Event(), Event()
        w0 = time.perf_counter()
        e0.record(stream)
        for arr, nj in one:
            ws.launch(arr, nj, st.flags, stream)
            stream.synchronize()
        e1.record(stream)
        stream.synchronize()
        return e0.elapsed_ms(e1), (time.perf_counter() - w0) * 1e3

    sweep = None
    if args.tune == "auto":
        tuning, table = autotune.tune(ctx, batched, log=log)
        sweep = [dict(c, ms=round(ms, 3)) for c, ms in table]
        log(f"autotune picked {tuning}")
    else:
        tuning = dict(checkpoint_k=args.checkpoint_k, pairs_per_lane=args.pairs_per_lane, schedule=args.schedule,
                      xcd_order=bool(args.xcd_order))
        ctx.set_tuning(**tuning)
    for _ in range(max(1, args.warmup)):
        batched()
    synchronize()
    times = [batched() for _ in range(args.steps)]
    ms = sum(times) / len(times)
    log(f"batched: {ms:.2f} ms/launch over {args.steps} launches")

    # Snapshot rows of every job from the batched run, then rerun plain
    # into the same buffers and compare bit for bit.
    def rows_of(k):
        j = jobs[k]
        P = int(st.pair_off[k + 1] - st.pair_off[k])
        out = []
        for r in sorted({0, j.n_sites // 2, j.n_sites - 1}):
            buf = np.empty(P, np.float32)
            _core.check(_core.lib().gsmc_memcpy_dtoh(ctypes.c_void_p(buf.ctypes.data),
                                                     st.d_out.offset(int(st.out_off[k]) + r * P), P * 4))
            out.append(buf)
        return np.concatenate(out)

    snap = [rows_of(k) for k in range(len(jobs))]
    plain_ev, plain_wall = plain()
    plain_ev2, plain_wall2 = plain()
    plain_ms = min(plain_ev, plain_ev2)
    log(f"plain: {plain_ms:.2f} ms (events), wall {min(plain_wall, plain_wall2):.2f} ms for {len(jobs)} launches")
    checks = {"batched_equals_plain_bit_exact": all(np.array_equal(snap[k].view(np.uint32), rows_of(k).view(np.uint32))
                                                    for k in range(len(jobs)))}

    cpu = None
    if not args.no_checks:
        from oracle import oracle as O

        T = O.tables_f32(DEFAULT_PATH, 1000, Ne, kmu, krho)
        # the longest job among those with at most 2,000 pairs (its whole
        # output comes back for the check), 512 of its pairs on the CPU
        k = max((i for i in range(len(jobs)) if len(jobs[i].pairs) <= 2000), key=lambda i: jobs[i].n_sites)
        j = jobs[k]
        m = min(512, len(j.pairs))
        sp = [tuple(map(int, p)) for p in j.pair_array()[:m]]
        t1 = time.perf_counter()
        ref = O.fb_blockwise_f32(np.ascontiguousarray(G[:, j.site_lo:j.site_hi]), pos[j.site_lo:j.site_hi], sp, T, Ne, wl["core"],
                                 wl["flank"])
        cpu_s = time.perf_counter() - t1
        P = len(j.pairs)
        full = np.empty((j.n_sites, P), np.float32)
        _core.check(_core.lib().gsmc_memcpy_dtoh(ctypes.c_void_p(full.ctypes.data), st.d_out.offset(int(st.out_off[k])),
                                                 full.nbytes))
        checks["oracle_bit_exact_sampled_job"] = bool(np.array_equal(ref["mean"].view(np.uint32),
                                                                      np.ascontiguousarray(full[:, :m]).view(np.uint32)))
        cores = int(os.environ.get("OMP_NUM_THREADS", "0")) or len(os.sched_getaffinity(0))
        cpu = {"value": m * j.n_sites / cpu_s, "unit": "pair-sites/s", "cores": min(cores, m), "kind": "port",
               "sample": f"{m} pairs x {j.n_sites} sites of job {k} (its blocks), {cpu_s:.2f}s"}

    K = tuning["checkpoint_k"]
    B = sum(len(j.pairs) * j.n_sites * (4.0 + 16.0 / K) + 8 * j.n_sites + 8 * len(j.pairs) for j in jobs) \
        + n * ((S + 63) // 64) * 8
    achieved = B / (ms / 1e3) / 1e9
    line = {
        "metric": "pair-site posteriors/sec (batched ragged jobs)",
        "value": pair_sites / (ms / 1e3),
        "unit": "pair-sites/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": wl["label"], "n_jobs": len(jobs), "pair_sites": pair_sites, "tuning": tuning,
                   "mode": "mean_only", "parallelism": "single GPU"},
        "plain": {"value": pair_sites / (plain_ms / 1e3), "ms": plain_ms, "launches": len(jobs)},
        "speedup_vs_plain": plain_ms / ms,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "fb_fwd + fb_bwd (one batched launch pair)", "algorithmic_bytes_per_launch": B},
        "cpu_baseline": cpu,
        "checks": checks,
        "autotune": sweep,
    }
    print(json.dumps(line), flush=True)
    if not all(checks.values()):
        sys.exit(1)


def run_serving(args, wl):
    """Serving loop: Poisson query streams at swept rates against the
    resident cohort, continuous batching vs one query per launch. Latency
    is arrival to device completion of the query's launch."""
    from gamma_smc_cu_amd import _core, serving, synthetic
    from gamma_smc_cu_amd.flow_field import DEFAULT_PATH
    from gamma_smc_cu_amd.infer import _estimate_scaled_params_packed

    n, S = wl["n"], wl["S"]
    t0 = time.perf_counter()
    packed, pos = synthetic.resident_cohort(n, S, wl["span"], wl["seed"])
    log(f"resident cohort {n} x {S} ({packed.nbytes / 1e6:.0f} MB packed) in {time.perf_counter() - t0:.1f}s")
    Ne = 1e4
    kmu, krho = _estimate_scaled_params_packed(packed, n, pos, 1.25e-8, 1e-8, Ne)
    _core.set_device(0)
    ctx = _core.FlowContext(packed, pos, Ne, kmu, krho, DEFAULT_PATH, 0, packed=True, n_sites=S)
    del packed
    ctx.set_tuning(checkpoint_k=args.checkpoint_k, pairs_per_lane=args.pairs_per_lane, schedule=args.schedule,
                   xcd_order=bool(args.xcd_order))
    arena = int(wl["arena_gb"] * 1e9)

    def one_run(mode, rate, nq):
        qs, off = serving.poisson_queries(nq, rate, n, S, wl["seed"], core_block_sites=wl["core"])
        if mode == "batched":
            srv = serving.Server(ctx, arena, n_slots=2, ck_bytes_per_launch=int(wl["ck_gb"] * 1e9))
        else:
            srv = serving.Server(ctx, arena, n_slots=1, ck_bytes_per_launch=int(wl["ck_gb"] * 1e9),
                                 max_queries_per_launch=1)
        start = serving.run_open_loop(srv, qs, off)
        lat = np.array([q.latency for q in qs]) * 1e3
        end = max(q.finished for q in qs)
        ps = float(sum(q.pair_sites for q in qs))
        r = {"mode": mode, "rate_per_s": rate, "queries": nq, "launches": srv.launches,
             "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
             "mean_ms": float(lat.mean()), "pair_sites": ps, "makespan_s": end - start,
             "posteriors_per_s": ps / (end - start), "max_in_flight": srv.max_in_flight,
             "arena_peak_gb": srv.arena.peak / 1e9}
        log(f"{mode:8s} rate {rate:7.1f}/s: p50 {r['p50_ms']:8.2f} ms p99 {r['p99_ms']:8.2f} ms "
            f"{r['posteriors_per_s']:.3e} pair-sites/s, {srv.launches} launches")
        del srv
        return r

    # warm both paths (first launches of a shape, page-in)
    one_run("batched", wl["rates"][0], 8)
    one_run("single", wl["rates"][0], 4)
    rows = []
    for rate in wl["rates"]:
        nq = int(min(wl["max_queries"], max(40, rate * wl["seconds"])))
        rows.append(one_run("batched", rate, nq))
        rows.append(one_run("single", rate, nq))

    # equality check: the same queries alone vs in the mix (bounded sample)
    qs, off = serving.poisson_queries(12, 1e6, n, S, wl["seed"] + 1, core_block_sites=wl["core"])
    qs = [q for q in qs if len(q.haplotypes) <= 60][:6] or qs[:2]
    mix = serving.Server(ctx, arena, n_slots=2, ck_bytes_per_launch=int(wl["ck_gb"] * 1e9))
    res = {}

    def keep(q):
        res[q.qid] = mix.result(q)
        mix.release(q)

    serving.run_open_loop(mix, qs, np.zeros(len(qs)), keep)
    alone = serving.Server(ctx, arena // 4, n_slots=1, ck_bytes_per_launch=int(wl["ck_gb"] * 1e9),
                           max_queries_per_launch=1)
    same = True
    for q in qs:
        c = serving.Query(q.haplotypes, q.site_lo, q.site_hi, q.core_block_sites, q.flank_sites)
        alone.submit(c)
        alone.drain()
        same &= bool(np.array_equal(alone.result(c).view(np.uint32), res[q.qid].view(np.uint32)))
        alone.release(c)
    checks = {"alone_vs_mix_bit_exact": same}

    top = max(wl["rates"])
    b = next(r for r in rows if r["mode"] == "batched" and r["rate_per_s"] == top)
    s1 = next(r for r in rows if r["mode"] == "single" and r["rate_per_s"] == top)
    line = {
        "metric": "serving pair-site posteriors/sec at the top offered rate (p50/p99 latency reported)",
        "value": b["posteriors_per_s"],
        "unit": "pair-sites/s",
        "n_gpus": 1,
        "steps": len(rows),
        "warmup": 2,
        "ms_per_step": None,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": wl["label"], "tuning": {"checkpoint_k": args.checkpoint_k,
                                                       "pairs_per_lane": args.pairs_per_lane,
                                                       "schedule": args.schedule},
                   "arena_gb": wl["arena_gb"], "ck_gb_per_launch": wl["ck_gb"], "slots": 2,
                   "latency": "query arrival to device completion of its launch",
                   "parallelism": "single GPU"},
        "p50_ms": b["p50_ms"],
        "p99_ms": b["p99_ms"],
        "single_query_per_launch": {"value": s1["posteriors_per_s"], "p50_ms": s1["p50_ms"], "p99_ms": s1["p99_ms"]},
        "speedup_vs_single": b["posteriors_per_s"] / s1["posteriors_per_s"],
        "sweep": rows,
        "roofline": None,
        "cpu_baseline": None,
        "checks": checks,
    }
    print(json.dumps(line), flush=True)
    if not all(checks.values()):
        sys.exit(1)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="chromosome", choices=sorted(WORKLOADS))
    ap.add_argument("--tune", default="auto", choices=("auto", "none"),
                    help="auto: sweep launch shapes first and keep the fastest (results are identical)")
    ap.add_argument("--checkpoint-k", type=int, default=16)
    ap.add_argument("--pairs-per-lane", type=int, default=1)
    ap.add_argument("--schedule", default="static", choices=("static", "persistent"))
    ap.add_argument("--xcd-order", type=int, default=0)
    ap.add_argument("--cpu-pairs", type=int, default=8000)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-checks", action="store_true",
                    help="skip the sampled-pairs checks (profiling runs: only same-shape decode launches)")
    ap.add_argument("--tuning-json", default=None,
                    help="take the launch shape from the config.tuning of a previous bench JSON line")
    ap.add_argument("--fp64-pairs", type=int, default=None,
                    help="pairs of the sample whose FP32 posterior is compared with the FP64 oracle")
    ap.add_argument("--graph-steps", type=int, default=2,
                    help="also time the decode captured into a HIP graph and replayed this many times")
    ap.add_argument("--codes-steps", type=int, default=2,
                    help="also time the decode with uint8 log-quantized output this many times")
    ap.add_argument("--fwd-steps", type=int, default=2,
                    help="also time forward-only filtering (run_fwd) this many times")
    ap.add_argument("--summary-steps", type=int, default=2,
                    help="also time the per-site summary mode (run_fb_summary) this many times")
    ap.add_argument("--no-host-stream", action="store_true",
                    help="skip the host-output (PCIe) run of the biobank workload")
    ap.add_argument("--copy-calibration-gib", type=float, default=1.0,
                    help="size of the plain copy kernel run after the timed region (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 "
                                   "(the decoder would shard over pairs with no communication)"}))
        sys.exit(2)

    wl = WORKLOADS[args.workload]
    if args.workload == "ragged":
        run_ragged(args, wl)
        return
    if args.workload == "serving":
        run_serving(args, wl)
        return

    from gamma_smc_cu_amd import _core, synthetic
    from gamma_smc_cu_amd.device import DeviceArray, Event, Stream, run_fb_device, synchronize
    from gamma_smc_cu_amd.flow_field import DEFAULT_PATH
    from gamma_smc_cu_amd.infer import _estimate_scaled_params, _recommend_core_block_size

    n, S = wl["n"], wl["S"]
    t0 = time.perf_counter()
    G, pos = synthetic.cohort(n, S, wl["span"], wl["seed"])
    log(f"cohort {n} x {S} generated in {time.perf_counter() - t0:.1f}s")
    Ne = 1e4
    kmu, krho = _estimate_scaled_params(G, pos, 1.25e-8, 1e-8, Ne)
    pairs = np.array([(i, j) for i in range(n) for j in range(i)], dtype=np.int32)
    P = pairs.shape[0]
    _core.set_device(0)
    free0, total0 = _core.cuda_mem_info()
    core, _ = _recommend_core_block_size(S, P, True, wl["flank"], free0)
    flank = wl["flank"]
    n_blocks = math.ceil(S / core)
    log(f"P={P} core={core} flank={flank} blocks={n_blocks} free={free0 / 1e9:.1f}GB")

    ctx = _core.FlowContext(G, pos, Ne, kmu, krho, DEFAULT_PATH, 0)
    d_pi = DeviceArray.from_host(pairs[:, 0].copy())
    d_pj = DeviceArray.from_host(pairs[:, 1].copy())
    d_out = DeviceArray((S, P), np.float32)
    stream = Stream()

    def run_once():
        e0, e1 = Event(), Event()
        e0.record(stream)
        run_fb_device(ctx, d_pi, d_pj, P, core, flank, 0, d_out, stream=stream)
        e1.record(stream)
        stream.synchronize()
        return e0.elapsed_ms(e1)

    tuning = dict(checkpoint_k=args.checkpoint_k, pairs_per_lane=args.pairs_per_lane, schedule=args.schedule,
                  xcd_order=bool(args.xcd_order))
    sweep = None
    if args.tuning_json:
        with open(args.tuning_json) as fh:
            tuning = json.loads(fh.read().strip().splitlines()[-1])["config"]["tuning"]
        ctx.set_tuning(**tuning)
    elif args.tune == "auto":
        from gamma_smc_cu_amd import autotune

        tuning, table = autotune.tune(ctx, run_once, log=log)
        sweep = [dict(c, ms=round(ms, 3)) for c, ms in table]
        log(f"autotune picked {tuning}")
    else:
        ctx.set_tuning(**tuning)

    for i in range(args.warmup):
        run_fb_device(ctx, d_pi, d_pj, P, core, flank, 0, d_out, stream=stream)
        stream.synchronize()
        log(f"warmup {i + 1}/{args.warmup} done")
    synchronize()
    ev0, ev1 = Event(), Event()
    w0 = time.perf_counter()
    ev0.record(stream)
    for i in range(args.steps):
        run_fb_device(ctx, d_pi, d_pj, P, core, flank, 0, d_out, stream=stream)
    ev1.record(stream)
    stream.synchronize()
    synchronize()
    wall = time.perf_counter() - w0
    ms = ev0.elapsed_ms(ev1) / args.steps
    log(f"timed {args.steps} steps: {ms:.2f} ms/step (events), wall {wall * 1e3 / args.steps:.2f} ms/step")

    pair_sites = float(P) * S
    value = pair_sites / (ms / 1e3)
    B = algorithmic_bytes(P, S, n, tuning["checkpoint_k"])

    # A plain copy kernel of known traffic, outside the timed region, so a
    # counter pass over this program can scale FETCH_SIZE / WRITE_SIZE.
    copy_bytes = int(args.copy_calibration_gib * (1 << 30)) // 16 * 16
    if copy_bytes:
        from gamma_smc_cu_amd._core import check as _check, lib as _lib

        src = DeviceArray((copy_bytes,), np.uint8)
        dst = DeviceArray((copy_bytes,), np.uint8)
        _check(_lib().gsmc_device_copy(dst.ptr, src.ptr, copy_bytes, stream.handle))
        stream.synchronize()
        del src, dst
    achieved = B / (ms / 1e3) / 1e9

    fwd_only = None
    if args.fwd_steps > 0:
        # Forward-only filtering (run_fwd) of all pairs over the full
        # sequence into the same device output; a separate, shorter
        # recursion (one FP16 table lookup per site, no checkpoints).
        from gamma_smc_cu_amd.device import run_fwd_device

        run_fwd_device(ctx, d_pi, d_pj, P, 0, d_out, stream=stream)
        stream.synchronize()
        e0, e1 = Event(), Event()
        e0.record(stream)
        for _ in range(args.fwd_steps):
            run_fwd_device(ctx, d_pi, d_pj, P, 0, d_out, stream=stream)
        e1.record(stream)
        stream.synchronize()
        fm = e0.elapsed_ms(e1) / args.fwd_steps
        fwd_only = {"ms": fm, "pair_sites_per_s": pair_sites / (fm / 1e3), "steps": args.fwd_steps,
                    "mode": "run_fwd, full sequence, mean only"}
        log(f"forward-only: {fm:.2f} ms/call")

    graph = None
    if args.graph_steps > 0:
        # The same decode captured once into a HIP graph and replayed.
        from gamma_smc_cu_amd._core import Job
        from gamma_smc_cu_amd.batch import Graph

        jb = (Job * 1)()
        jb[0].site_lo, jb[0].site_hi, jb[0].core_block_sites, jb[0].flank_sites = 0, S, core, flank
        jb[0].pair_i, jb[0].pair_j, jb[0].n_pairs, jb[0].mean = d_pi.ptr.value, d_pj.ptr.value, P, d_out.ptr.value
        gr = Graph(ctx, jb, 1, 0)
        gr.launch(stream)
        stream.synchronize()
        e0, e1 = Event(), Event()
        e0.record(stream)
        for _ in range(args.graph_steps):
            gr.launch(stream)
        e1.record(stream)
        stream.synchronize()
        gm = e0.elapsed_ms(e1) / args.graph_steps
        graph = {"ms": gm, "pair_sites_per_s": pair_sites / (gm / 1e3), "steps": args.graph_steps,
                 "mode": "captured decode replayed from a HIP graph"}
        log(f"graph replay: {gm:.2f} ms/step")
        gr.close()

    codes = None
    if args.codes_steps > 0:
        # The same decode with the mean written as uint8 log codes (1 B per
        # pair-site instead of 4).
        from gamma_smc_cu_amd._core import GSMC_Q8

        d_q = DeviceArray((S, P), np.uint8)
        run_fb_device(ctx, d_pi, d_pj, P, core, flank, GSMC_Q8, d_q, stream=stream)
        stream.synchronize()
        e0, e1 = Event(), Event()
        e0.record(stream)
        for _ in range(args.codes_steps):
            run_fb_device(ctx, d_pi, d_pj, P, core, flank, GSMC_Q8, d_q, stream=stream)
        e1.record(stream)
        stream.synchronize()
        cm = e0.elapsed_ms(e1) / args.codes_steps
        codes = {"ms": cm, "pair_sites_per_s": pair_sites / (cm / 1e3), "steps": args.codes_steps,
                 "output_bytes": P * S, "mode": "uint8 log-quantized mean codes"}
        log(f"uint8 codes: {cm:.2f} ms/step")
        del d_q

    summary = None
    if args.summary_steps > 0:
        # Fused per-site reduction over all pairs (no per-pair output): the
        # same decode with a 16 B per 64-pair group and site epilogue.
        ctx.run_fb_summary(pairs, core_block_sites=core, flank_sites=flank)
        t1 = time.perf_counter()
        for _ in range(args.summary_steps):
            ctx.run_fb_summary(pairs, core_block_sites=core, flank_sites=flank)
        sm = (time.perf_counter() - t1) * 1e3 / args.summary_steps
        summary = {"ms": sm, "pair_sites_per_s": pair_sites / (sm / 1e3), "steps": args.summary_steps,
                   "timing": "host wall clock per call (pair upload, decode, 3 x n_sites floats back)"}
        log(f"summary mode: {sm:.2f} ms/call")

    checks = {}
    cpu = fp64 = host = None
    if args.fp64_pairs is None:
        args.fp64_pairs = wl["fp64_pairs"]
    if not args.no_checks:
        checks, cpu, fp64, (sample, alone) = sample_checks(args, ctx, G, pos, pairs, P, S, core, flank, Ne, kmu,
                                                           krho, d_out, stream)
        if wl.get("host_stream") and not args.no_host_stream:
            del d_out
            host = host_stream_run(ctx, pairs, core, flank, sample, alone)
            checks["host_stream_matches_device"] = host.pop("sample_matches_device_run")

    line = {
        "metric": "pair-site posteriors/sec",
        "value": value,
        "unit": "pair-sites/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": wl["label"], "n_haplotypes": n, "n_pairs": P, "n_sites": S,
                   "core_block_sites": core, "flank_sites": flank, "n_blocks": n_blocks,
                   "tuning": tuning, "mode": "mean_only", "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": MEASURED_TRAFFIC.get((args.workload, tuning["checkpoint_k"],
                                                      tuning["pairs_per_lane"], tuning["schedule"],
                                                      bool(tuning["xcd_order"]))),
                     "kernel": ("fb_fwd_items + fb_bwd_items" if tuning["schedule"] == "static"
                                else "fb_fwd_queue + fb_bwd_queue"),
                     "algorithmic_bytes_per_launch": B},
        "cpu_baseline": cpu,
        "fp64_error": fp64,
        "host_stream": host,
        "site_summary": summary,
        "forward_only": fwd_only,
        "codes_u8": codes,
        "graph_replay": graph,
        "checks": checks,
        "wall_ms_per_step": wall * 1e3 / args.steps,
        "copy_calibration_bytes": copy_bytes,
        "autotune": sweep,
    }
    print(json.dumps(line), flush=True)
    if not all(checks.values()):
        sys.exit(1)


if __name__ == "__main__":
    main()
