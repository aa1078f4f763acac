"""Benchmark: pair-site posteriors/s of the blockwise decoder on one MI355X.

Default workload (``chromosome``): 356 haplotypes (all 63,190 pairs) x
402,853 segregating sites, positions uniform on [0, 51 Mb), 1/k spectrum,
seed 7; ``infer_blockwise`` mean_only semantics with 2048-site flanks and
the core block size the auto-sizer picks from free device memory. One step
is one decode of every pair over every block, with the packed cohort, pair
ids and output resident in HBM when the timed region starts; ``value`` is
pair-sites per second of the timed steps (HIP events on the decode stream).

The one JSON line also carries:
* ``roofline``: the decode's algorithmic HBM bytes per launch over its
  event-timed duration against the 8 TB/s peak (``frac``), the same against
  the 4 B per pair-site output floor alone (``frac_output_floor``), measured
  HBM traffic when the counter summary matches these kernels, and the
  binding unit, the texture address path: the decode's table lookups per
  second against what a gather-only probe kernel sustains on the same
  tables in this run (``ta``);
* ``cpu_baseline``: the CPU oracle (FP32 restatement of the same
  recursion, OpenMP over pairs) on this host: a fixed 1,000-pair random
  sample over all sites, one warm-up then the best of 3, linearly
  extrapolated to all pairs, with the CPU model and thread counts;
* ``checks``: the sampled pairs decoded alone on the GPU equal the oracle
  bit for bit and the same columns of the full run; FP64 error bounds;
* ``end_to_end``: ``infer_blockwise`` as the reference's headline run calls
  it (all pairs, auto theta, host output arrays), wall time beside the
  reference's published 23.659 s;
* ``biobank``, ``ragged``, ``serving``: compact runs of the other workloads
  (``--workload`` runs any of them alone, at full length), each with its
  own checks;
* ``gene_stats``: the per-gene accumulators of the reference's production
  pipeline for all pairs over 800 synthetic genes, computed on the device
  (only the accumulators come back), checked on sampled pairs;
* ``plain``: the reference's small one-shot call (190 pairs x 17,339
  sites, ``gamma_smc_flow_cached_fb``) beside its published 0.0333 s;
* ``scaling``: ``infer()`` over all pairs at the reference's pairwise
  scaling shapes (4,950 / 19,900 / 79,800 pairs) beside its published
  times.
"""

from __future__ import annotations

import argparse
import ctypes
import gc
import hashlib
import json
import math
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
REFERENCE_E2E_S = 23.659  # benchmarks/pairwise_scaling/results.csv:22 (1x B200, all 63,190 pairs)
# results.csv:17-21: the same call on the first k shuffled pairs
REFERENCE_E2E_SUBSETS = [(1, 1.1561), (10, 1.2289), (100, 1.3497), (1000, 1.6851), (10000, 4.6351)]

# HBM bytes per decode launch pair measured with rocprofv3 FETCH_SIZE and
# WRITE_SIZE passes (each scaled by what the counter reads for the
# calibration copy kernel), per workload and launch shape, valid for the
# kernel source with this SHA-256 (gamma_smc_cu_amd/csrc/fb_kernels.h): the
# counter summary under profiles/ gives the commands. Any other kernel
# source reports traffic null.
MEASURED_TRAFFIC = {
    "kernel_sha256": "5f74bb52d1d56b2fc3203d81c7af9bc329c0c5199468dd396a29812a0a84e91c",
    "entries": {
        "chromosome/K8/np1/dynamic/xcd0": 195.32e9,
        "chromosome/K8/np1/persistent/xcd0": 189.00e9,
        "chromosome/K8/np1/static/xcd0": 197.57e9,
    },
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


def block_windows(S, core, flank):
    return [(cs, min(S, cs + core), max(0, cs - flank), min(S, cs + core + flank)) for cs in range(0, S, core)]


def decode_lookups(P, S, core, flank, K):
    """Table lookups of one decode: per block, the forward walk over
    [ps, ce) and the backward walk over the right flank (pe - ps in all),
    then per core site the backward step and, with checkpoints, the
    recomputed forward step of every site but the first of each K-site
    window, whose state is the checkpoint (stored forward states, K = 1,
    recompute none). Gaps longer than the table depth add stretch lookups;
    the workloads here have none."""
    per = 0
    for cs, ce, ps, pe in block_windows(S, core, flank):
        n = ce - cs
        recompute = 0 if K == 1 else n - (n + K - 1) // K
        per += (pe - ps) + n + recompute
    return P * per


def kernel_sha256():
    with open(os.path.join(ROOT, "gamma_smc_cu_amd", "csrc", "fb_kernels.h"), "rb") as fh:
        return hashlib.sha256(fh.read()).hexdigest()


def measured_traffic(workload, tuning):
    if MEASURED_TRAFFIC["kernel_sha256"] != kernel_sha256():
        return None
    key = f"{workload}/K{tuning['checkpoint_k']}/np{tuning['pairs_per_lane']}/{tuning['schedule']}" \
          f"/xcd{int(bool(tuning['xcd_order']))}"
    return MEASURED_TRAFFIC["entries"].get(key)


def cpu_info():
    model = platform.processor() or "unknown"
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    affinity = len(os.sched_getaffinity(0))
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or affinity
    return {"cpu_model": model, "nproc": os.cpu_count(), "affinity_cpus": affinity, "omp_threads": threads}


def probe_ta(ctx, stream):
    """Lookups per second of the gather-only probe on this context's tables."""
    from gamma_smc_cu_amd._core import check, lib
    from gamma_smc_cu_amd.device import DeviceArray

    waves, steps = 8192, 2048
    out = DeviceArray((waves * 64,), np.float32)
    ms = ctypes.c_float()
    check(lib().gsmc_lookup_probe(ctx.handle, waves, steps, out.ptr, stream.handle, ctypes.byref(ms)))
    return 2.0 * 64 * waves * steps / (ms.value / 1e3)


def rows_of(d_out, rows, P, cols=None):
    from gamma_smc_cu_amd._core import check, lib

    n = P if cols is None else cols
    out = np.empty((len(rows), n), np.float32)
    for k, r in enumerate(rows):
        check(lib().gsmc_memcpy_dtoh(ctypes.c_void_p(out[k].ctypes.data), d_out.offset(int(r) * P), n * 4))
    return out


def bits_equal(a, b):
    a, b = np.ascontiguousarray(a), np.ascontiguousarray(b)
    return bool(a.shape == b.shape and np.array_equal(a.view(np.uint32), b.view(np.uint32)))


def decode_sample_alone(ctx, pairs, sample, S, core, flank, stream):
    from gamma_smc_cu_amd.device import DeviceArray, run_fb_device

    sp = pairs[sample]
    d_spi = DeviceArray.from_host(np.ascontiguousarray(sp[:, 0]))
    d_spj = DeviceArray.from_host(np.ascontiguousarray(sp[:, 1]))
    d_sout = DeviceArray((S, len(sample)), np.float32)
    run_fb_device(ctx, d_spi, d_spj, len(sample), core, flank, 0, d_sout, stream=stream)
    stream.synchronize()
    return d_sout.to_host()


def cpu_baseline_and_checks(G, pos, pairs, P, S, core, flank, Ne, kmu, krho, alone, sample, fp64_pairs,
                            cpu_pairs, no_cpu):
    """The CPU oracle on the sampled pairs: one warm-up, best of 3 (timed),
    its output compared bit for bit with the pairs decoded alone on the GPU;
    FP64 error on the first fp64_pairs of them."""
    from gamma_smc_cu_amd.flow_field import DEFAULT_PATH
    from oracle import oracle as O

    checks, cpu, fp64 = {}, None, None
    sp = [tuple(map(int, pairs[k])) for k in sample]
    T = O.tables_f32(DEFAULT_PATH, 1000, Ne, kmu, krho)
    m = len(sample) if no_cpu else min(cpu_pairs, len(sample))
    ref = O.fb_blockwise_f32(G, pos, sp[:m], T, Ne, core, flank)  # warm-up run, also the parity check
    checks["oracle_bit_exact_sample"] = bits_equal(ref["mean"], alone[:, :m])
    if not no_cpu:
        times = []
        for _ in range(3):
            t1 = time.perf_counter()
            O.fb_blockwise_f32(G, pos, sp[:m], T, Ne, core, flank)
            times.append(time.perf_counter() - t1)
        best = min(times)
        info = cpu_info()
        cpu = {"value": m * S / best, "unit": "pair-sites/s", "cores": info["omp_threads"], "kind": "port",
               "sample": f"{m} random pairs (seed 123) x all {S} sites, same blocks; one warm-up, best of 3 "
                         f"({', '.join(f'{t:.2f}' for t in times)} s); all {P} pairs extrapolate linearly "
                         f"to {best * P / m:.0f} s",
               **info}
        log(f"cpu baseline: {cpu['value']:.3e} pair-sites/s on {info['omp_threads']} threads "
            f"({info['cpu_model']}, best {best:.2f}s)")
    k = min(fp64_pairs, len(sample))
    if k > 0:
        T64 = O.tables_f64(DEFAULT_PATH, 1000, Ne, kmu, krho)
        t1 = time.perf_counter()
        ref64 = O.fb_blockwise_f64(G, pos, sp[:k], T64, Ne, core, flank)
        fp64 = O.log_error_report(alone[:, :k], ref64["mean"])
        fp64["sample"] = f"{k} random pairs x all {S} sites"
        fp64["tolerance"] = "max |log mean32 - log mean64| <= 1e-2, median <= 1e-3"
        checks["fp64_error_within_tolerance"] = bool(fp64["max"] <= 1e-2 and fp64["median"] <= 1e-3)
        log(f"fp64 error: max {fp64['max']:.2e} median {fp64['median']:.2e} ({time.perf_counter() - t1:.1f}s)")
    return checks, cpu, fp64


REFERENCE_PLAIN_S = 0.0333  # benchmarks/test_suite_stdpopsim/results/config_000.json (B200 MIG slice)


def plain_small():
    """The reference's small-job figure (run_one.py:244-260): a one-shot
    ``_core.gamma_smc_flow_cached_fb`` over all 190 pairs of 20 haplotypes,
    mean only, auto theta, min of 3 after a warm-up. Its input there is a
    5 Mb msprime simulation with 17,339 sites (config_000.json); here a 1/k
    cohort of the same shape (seed 1). Checked bit for bit against the
    oracle."""
    from gamma_smc_cu_amd import _core, synthetic
    from gamma_smc_cu_amd.flow_field import DEFAULT_PATH
    from gamma_smc_cu_amd.infer import _estimate_scaled_params
    from oracle import oracle as O

    n, S, span, Ne = 20, 17_339, 5_000_000, 1e4
    G, pos = synthetic.cohort(n, S, span, 1)
    pairs = [(i, j) for i in range(n) for j in range(i + 1, n)]
    kmu, krho = _estimate_scaled_params(G, pos, 1.29e-8, 1.2820402396300887e-8, Ne)
    _core.gamma_smc_flow_cached_fb(G, pos, [(0, 1)], Ne, kmu, krho, DEFAULT_PATH, True, 0)  # warm-up
    times, mean = [], None
    for _ in range(3):
        t0 = time.perf_counter()
        out = _core.gamma_smc_flow_cached_fb(G, pos, pairs, Ne, kmu, krho, DEFAULT_PATH, True, 0)["mean"]
        times.append(time.perf_counter() - t0)
        mean = out if mean is None else mean
    T = O.tables_f32(DEFAULT_PATH, 1000, Ne, kmu, krho)
    ref = O.fb_window_f32(G, pos, pairs, T, Ne)["mean"]
    t = min(times)
    out = {"seconds": t, "all_seconds": times, "pair_sites_per_s": len(pairs) * S / t,
           "reference_seconds": REFERENCE_PLAIN_S, "speedup_vs_reference": REFERENCE_PLAIN_S / t,
           "call": "_core.gamma_smc_flow_cached_fb(G, pos, all 190 pairs, Ne, auto theta mu, rho, flow field, "
                   "mean_only=True, 0), one-shot (context, decode, host output), min of 3 after a warm-up",
           "input": f"{n} haplotypes x {S} sites over 5 Mb, 1/k spectrum, seed 1 (the reference: msprime, "
                    "HomSap OutOfAfrica_3G09, 17,339 sites)",
           "checks": {"oracle_bit_exact": bits_equal(mean, ref)}}
    log(f"plain: gamma_smc_flow_cached_fb {1e3 * t:.1f} ms for {len(pairs)} pairs x {S} sites "
        f"(reference {1e3 * REFERENCE_PLAIN_S:.1f} ms)")
    return out


# benchmarks/pairwise_scaling/scaling_results.csv:2-4 (1x B200): haplotypes, sites, seconds
REFERENCE_SCALING = [(100, 99_501, 0.806373119354248), (200, 117_259, 4.169840574264526),
                     (400, 130_296, 457.0374381542206)]


def scaling_infer():
    """``infer()`` over every pair as the reference's pairwise scaling run
    calls it (run_stdpopsim_scaling.py:58-76): a one-pair warm-up, then all
    pairs, mean only, auto theta, output into host arrays; 1/k cohorts of the
    published shapes over 51 Mb (seed 19) in place of its stdpopsim chr22
    simulations. Three columns of each checked bit for bit against the
    oracle."""
    import gamma_smc_cu_amd
    from gamma_smc_cu_amd import synthetic
    from gamma_smc_cu_amd.flow_field import DEFAULT_PATH
    from gamma_smc_cu_amd.infer import _estimate_scaled_params
    from oracle import oracle as O

    out = []
    for n, S, ref_s in REFERENCE_SCALING:
        G, pos = synthetic.cohort(n, S, 51_000_000, 19)
        gamma_smc_cu_amd.infer(G, pos, pairs=[(1, 0)], mean_only=True, auto_estimate_theta=True)
        t0 = time.perf_counter()
        r = gamma_smc_cu_amd.infer(G, pos, mean_only=True, auto_estimate_theta=True)
        t = time.perf_counter() - t0
        all_pairs = [(i, j) for i in range(n) for j in range(i)]
        P = len(all_pairs)
        cols = [0, P // 2, P - 1]
        kmu, krho = _estimate_scaled_params(G, pos, 1.25e-8, 1e-8, 1e4)
        ref = O.fb_window_f32(G, pos, [all_pairs[k] for k in cols], O.tables_f32(DEFAULT_PATH, 1000, 1e4, kmu, krho),
                              1e4)["mean"]
        ok = bits_equal(np.ascontiguousarray(r["mean"][:, cols]), ref)
        out.append({"n_haplotypes": n, "n_pairs": P, "n_sites": S, "seconds": t, "pair_sites_per_s": P * S / t,
                     "output_gb": 4.0 * P * S / 1e9, "reference_seconds": ref_s,
                     "speedup_vs_reference": ref_s / t, "checks": {"oracle_bit_exact_3_columns": ok}})
        log(f"scaling: infer() {n} haplotypes ({P} pairs) x {S} sites: {t:.2f}s (reference {ref_s:.2f}s)")
        del r, G, pos
        gc.collect()
    return {"call": "infer(G, pos, mu=1.25e-8, rho=1e-8, Ne=10000, mean_only=True, auto_estimate_theta=True), "
                    "all pairs, after a one-pair warm-up; (S, P) float32 host output",
            "input": "1/k-spectrum cohorts of the published shapes over 51 Mb, seed 19 (the reference: stdpopsim "
                     "chr22 simulations)",
            "reference_hardware": "1x NVIDIA B200 (benchmarks/pairwise_scaling/scaling_results.csv:2-4)",
            "runs": out,
            "checks": {f"n{r['n_haplotypes']}_oracle_bit_exact": r["checks"]["oracle_bit_exact_3_columns"]
                       for r in out}}


def end_to_end(G, pos, sample_pairs, alone, rows):
    """infer_blockwise as the reference's headline benchmark calls it
    (run_final_benchmark.py:51-63): a one-pair warm-up, then all pairs
    (i < j, shuffled with default_rng(42)), auto theta, pair_batch_size
    30,000, output into host arrays. Wall time of the second call."""
    import gamma_smc_cu_amd

    n = G.shape[0]
    all_pairs = [(i, j) for i in range(n) for j in range(i + 1, n)]
    np.random.default_rng(42).shuffle(all_pairs)
    gamma_smc_cu_amd.infer_blockwise(G, pos, mu=1.25e-8, rho=1e-8, Ne=10000, pairs=[(0, 1)], mean_only=True,
                                     auto_estimate_theta=True)
    subsets = []
    for k, ref_s in REFERENCE_E2E_SUBSETS:
        t0 = time.perf_counter()
        r = gamma_smc_cu_amd.infer_blockwise(G, pos, mu=1.25e-8, rho=1e-8, Ne=10000, pairs=all_pairs[:k],
                                             mean_only=True, auto_estimate_theta=True, pair_batch_size=min(k, 30000))
        subsets.append({"n_pairs": k, "seconds": time.perf_counter() - t0, "reference_seconds": ref_s})
        del r
    t0 = time.perf_counter()
    r = gamma_smc_cu_amd.infer_blockwise(G, pos, mu=1.25e-8, rho=1e-8, Ne=10000, pairs=all_pairs, mean_only=True,
                                         auto_estimate_theta=True, pair_batch_size=min(len(all_pairs), 30000))
    wall = time.perf_counter() - t0
    col = {p: k for k, p in enumerate(all_pairs)}
    cols = [col[(int(j), int(i))] for i, j in sample_pairs]  # the sample holds (i > j)
    got = r["mean"][np.asarray(rows)][:, cols]
    P, S = len(all_pairs), G.shape[1]
    out = {"seconds": wall, "pair_sites_per_s": P * S / wall, "output_gb": 4.0 * P * S / 1e9,
           "reference_seconds": REFERENCE_E2E_S, "speedup_vs_reference": REFERENCE_E2E_S / wall,
           "reference_hardware": "1x NVIDIA B200 (benchmarks/pairwise_scaling/results.csv:22)",
           "call": "infer_blockwise(G, pos, pairs=all pairs shuffled, mean_only=True, auto_estimate_theta=True, "
                   "pair_batch_size=30000), after a one-pair warm-up; G and positions in host memory, "
                   "output (S, P) float32 host array",
           "checks": {"sampled_columns_equal_device_run": bits_equal(got, alone[np.asarray(rows)])},
           "pair_subsets": subsets}
    log("end to end, pair subsets: " + ", ".join(f"{x['n_pairs']}: {x['seconds']:.3f}s (reference {x['reference_seconds']}s)"
                                                 for x in subsets))
    log(f"end to end: infer_blockwise {wall:.2f}s for {P} pairs x {S} sites "
        f"({out['output_gb']:.1f} GB to host; reference {REFERENCE_E2E_S}s)")
    del r
    gc.collect()
    return out


def run_decode_workload(args, name, compact=False):
    """The chromosome or biobank decode: headline numbers and checks."""
    from gamma_smc_cu_amd import _core, synthetic
    from gamma_smc_cu_amd.device import DeviceArray, Event, Stream, run_fb_device, synchronize
    from gamma_smc_cu_amd.flow_field import DEFAULT_PATH
    from gamma_smc_cu_amd.infer import _estimate_scaled_params, _recommend_core_block_size

    wl = WORKLOADS[name]
    n, S = wl["n"], wl["S"]
    t0 = time.perf_counter()
    G, pos = synthetic.cohort(n, S, wl["span"], wl["seed"])
    log(f"{name}: cohort {n} x {S} generated in {time.perf_counter() - t0:.1f}s")
    Ne = 1e4
    kmu, krho = _estimate_scaled_params(G, pos, 1.25e-8, 1e-8, Ne)
    a, b = np.tril_indices(n, k=-1)
    pairs = np.stack([a, b], axis=1).astype(np.int32)  # the all-pairs order (1,0), (2,0), (2,1), ...
    P = pairs.shape[0]
    _core.set_device(0)
    free0, _ = _core.cuda_mem_info()
    core, _ = _recommend_core_block_size(S, P, True, wl["flank"], free0)
    flank = wl["flank"]
    n_blocks = math.ceil(S / core)
    log(f"{name}: P={P} core={core} flank={flank} blocks={n_blocks} free={free0 / 1e9:.1f}GB")

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
        log(f"{name}: autotune picked {tuning}")
    else:
        ctx.set_tuning(**tuning)

    warmup, steps = (2, 3) if compact else (args.warmup, args.steps)
    for i in range(warmup):
        run_fb_device(ctx, d_pi, d_pj, P, core, flank, 0, d_out, stream=stream)
        stream.synchronize()
    synchronize()
    ev0, ev1 = Event(), Event()
    w0 = time.perf_counter()
    ev0.record(stream)
    for i in range(steps):
        run_fb_device(ctx, d_pi, d_pj, P, core, flank, 0, d_out, stream=stream)
    ev1.record(stream)
    stream.synchronize()
    synchronize()
    wall = time.perf_counter() - w0
    ms = ev0.elapsed_ms(ev1) / steps
    log(f"{name}: timed {steps} steps: {ms:.2f} ms/step (events), wall {wall * 1e3 / steps:.2f} ms/step")

    pair_sites = float(P) * S
    value = pair_sites / (ms / 1e3)
    K = tuning["checkpoint_k"]
    B = algorithmic_bytes(P, S, n, K)
    achieved = B / (ms / 1e3) / 1e9
    floor_bytes = 4.0 * P * S
    lookups = decode_lookups(P, S, core, flank, K)
    ta_peak = probe_ta(ctx, stream)
    ta_rate = lookups / (ms / 1e3)
    log(f"{name}: {lookups / pair_sites:.3f} lookups per pair-site, {ta_rate:.3e} lookups/s against the probe's "
        f"{ta_peak:.3e} ({ta_rate / ta_peak:.2f})")
    roofline = {
        "bound": "ta",
        "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
        "traffic": measured_traffic(name, tuning),
        "kernel": {"static": "fb_fwd_items + fb_bwd_items", "persistent": "fb_fwd_queue + fb_bwd_queue",
                   "dynamic": "fb_fwd_grab + fb_bwd_grab"}[tuning["schedule"]] if K != 1 else "stored-state kernels",
        "algorithmic_bytes_per_launch": B,
        "frac_output_floor": floor_bytes / (ms / 1e3) / 1e9 / HBM_PEAK_GBS,
        "output_floor_bytes_per_launch": floor_bytes,
        "ta": {"unit": "pair-lookups/s", "achieved": ta_rate, "peak": ta_peak, "frac": ta_rate / ta_peak,
               "lookups_per_launch": lookups, "lookups_per_pair_site": lookups / pair_sites,
               "peak_source": "gsmc_lookup_probe: gather-only chains of the decode's lookup on these tables, "
                              "2 per lane, 8 waves per SIMD, timed in this run"},
        "note": "bound names the unit the counters show saturated (texture address path, TA busy ~90%); "
                "achieved/frac are HBM algorithmic bytes against 8 TB/s; frac_output_floor counts only the "
                "4 B/pair-site posterior output, so a smaller checkpoint interval cannot raise it",
    }

    # A plain copy kernel of known traffic, outside the timed region, so a
    # counter pass over this program can scale FETCH_SIZE / WRITE_SIZE.
    copy_bytes = int(args.copy_calibration_gib * (1 << 30)) // 16 * 16 if not compact else 0
    if copy_bytes:
        src = DeviceArray((copy_bytes,), np.uint8)
        dst = DeviceArray((copy_bytes,), np.uint8)
        _core.check(_core.lib().gsmc_device_copy(dst.ptr, src.ptr, copy_bytes, stream.handle))
        stream.synchronize()
        del src, dst

    extras = {}
    if not compact:
        extras = timed_extras(args, ctx, d_pi, d_pj, d_out, P, S, core, flank, pairs, stream, pair_sites)

    out = {"name": name, "label": wl["label"], "P": P, "S": S, "n": n, "core": core, "flank": flank,
           "n_blocks": n_blocks, "tuning": tuning, "ms": ms, "value": value, "wall_ms_per_step": wall * 1e3 / steps,
           "steps": steps, "warmup": warmup, "roofline": roofline, "autotune": sweep, "copy_bytes": copy_bytes,
           **extras}
    if args.no_checks:
        out["checks"] = {}
        return out, None

    # sampled pairs alone vs the full run and the oracle
    rng = np.random.default_rng(123)
    m = min(max(args.cpu_pairs, wl["fp64_pairs"]), P)
    sample = np.sort(rng.choice(P, size=m, replace=False))
    alone = decode_sample_alone(ctx, pairs, sample, S, core, flank, stream)
    rows = np.unique(np.concatenate([[0, core - 1, min(core, S - 1), S // 2, S - 1],
                                     np.random.default_rng(5).integers(0, S, 24)]))
    full_rows = rows_of(d_out, rows, P)
    checks = {"alone_vs_mix_bit_exact": bits_equal(full_rows[:, sample], alone[rows]),
              "finite_positive": bool(np.isfinite(full_rows).all() and (full_rows > 0).all())}
    d_out.free()
    del d_out
    gc.collect()
    c2, cpu, fp64 = cpu_baseline_and_checks(G, pos, pairs, P, S, core, flank, Ne, kmu, krho, alone, sample,
                                            wl["fp64_pairs"], args.cpu_pairs, args.no_cpu_baseline or compact)
    checks.update(c2)
    out.update({"checks": checks, "cpu_baseline": cpu, "fp64_error": fp64})
    if not compact and args.gene_steps > 0:
        out["gene_stats"] = gene_stats_run(ctx, pairs, pos, core, flank, sample, alone, args.gene_steps, wl["span"])
        checks.update({f"gene_stats_{k}": v for k, v in out["gene_stats"]["checks"].items()})
    state = {"G": G, "pos": pos, "ctx": ctx, "pairs": pairs, "sample": sample, "alone": alone, "rows": rows}
    return out, state


def gene_stats_run(ctx, pairs, pos, core, flank, sample, alone, steps, span):
    """Per-gene statistics of all pairs (``run_fb_gene_stats``: the decode
    with the gene epilogue in place of row stores, then the per-gene
    reduction; only the accumulators come back) over 800 synthetic genes,
    host wall clock per call. Checked on the first 1,000 sampled pairs
    against the oracle's restatement of the reference's host accumulation
    over their device means."""
    from gamma_smc_cu_amd import synthetic
    from gamma_smc_cu_amd.genes import gene_site_ranges
    from oracle import oracle as O

    genes = synthetic.genes(span)
    ranges = gene_site_ranges(pos, genes)
    P, S = pairs.shape[0], pos.shape[0]
    ctx.run_fb_gene_stats(pairs, ranges, core_block_sites=core, flank_sites=flank)
    t0 = time.perf_counter()
    for _ in range(steps):
        acc = ctx.run_fb_gene_stats(pairs, ranges, core_block_sites=core, flank_sites=flank)
    ms = (time.perf_counter() - t0) * 1e3 / steps
    covered = np.zeros(S + 1, np.int64)
    np.add.at(covered, ranges[:, 0], 1)
    np.add.at(covered, ranges[:, 1], -1)
    n_cov = int(np.count_nonzero(np.cumsum(covered)[:S] > 0))
    k = min(1000, len(sample))
    got = ctx.run_fb_gene_stats(pairs[sample[:k]], ranges, core_block_sites=core, flank_sites=flank)
    want = O.gene_accumulators(alone[:, :k], pos, genes)
    sums = all(np.allclose(got[x], want[x], rtol=1e-5, atol=1e-6) for x in ("lin_sum", "log_sum", "log_sq_sum"))
    mins = all(np.allclose(got[x], want[x], rtol=1e-5, atol=1e-6) for x in ("min_lin", "min_log"))
    hd = int(np.abs(got["histogram"] - want["histogram"]).sum())
    checks = {"sample_counts_equal_oracle": bool(np.array_equal(got["count"], want["count"])),
              "sample_sums_and_minima_within_1e-5": bool(sums and mins),
              "sample_histogram_moves_le_0.1pct": bool(hd <= 1e-3 * max(int(want["histogram"].sum()), 1))}
    log(f"gene stats: {ms:.1f} ms/call over {len(genes)} genes ({n_cov} of {S} sites covered), "
        f"sample histogram moves {hd}")
    return {"ms": ms, "pair_sites_per_s": float(P) * S / (ms / 1e3), "steps": steps, "n_genes": len(genes),
            "covered_sites": n_cov, "histogram_bins": int(acc["histogram"].shape[1]),
            "timing": "host wall clock per call (pair upload, decode with the gene epilogue, per-gene "
                      "reduction, accumulators back)",
            "checks": checks}


def timed_extras(args, ctx, d_pi, d_pj, d_out, P, S, core, flank, pairs, stream, pair_sites):
    """Forward-only, graph replay, uint8 codes and per-site summary modes of
    the same decode (each a couple of calls)."""
    from gamma_smc_cu_amd.device import DeviceArray, Event, run_fb_device, run_fwd_device

    def timed(fn, n):
        fn()
        stream.synchronize()
        e0, e1 = Event(), Event()
        e0.record(stream)
        for _ in range(n):
            fn()
        e1.record(stream)
        stream.synchronize()
        return e0.elapsed_ms(e1) / n

    ex = {}
    if args.fwd_steps > 0:
        fm = timed(lambda: run_fwd_device(ctx, d_pi, d_pj, P, 0, d_out, stream=stream), args.fwd_steps)
        ex["forward_only"] = {"ms": fm, "pair_sites_per_s": pair_sites / (fm / 1e3), "steps": args.fwd_steps,
                              "mode": "run_fwd, full sequence, mean only"}
        log(f"forward-only: {fm:.2f} ms/call")
    if args.graph_steps > 0:
        from gamma_smc_cu_amd._core import Job
        from gamma_smc_cu_amd.batch import Graph

        jb = (Job * 1)()
        jb[0].site_lo, jb[0].site_hi, jb[0].core_block_sites, jb[0].flank_sites = 0, S, core, flank
        jb[0].pair_i, jb[0].pair_j, jb[0].n_pairs, jb[0].mean = d_pi.ptr.value, d_pj.ptr.value, P, d_out.ptr.value
        gr = Graph(ctx, jb, 1, 0)
        gm = timed(lambda: gr.launch(stream), args.graph_steps)
        ex["graph_replay"] = {"ms": gm, "pair_sites_per_s": pair_sites / (gm / 1e3), "steps": args.graph_steps,
                              "mode": "captured decode replayed from a HIP graph"}
        log(f"graph replay: {gm:.2f} ms/step")
        gr.close()
    if args.codes_steps > 0:
        from gamma_smc_cu_amd._core import GSMC_Q8

        d_q = DeviceArray((S, P), np.uint8)
        cm = timed(lambda: run_fb_device(ctx, d_pi, d_pj, P, core, flank, GSMC_Q8, d_q, stream=stream),
                   args.codes_steps)
        ex["codes_u8"] = {"ms": cm, "pair_sites_per_s": pair_sites / (cm / 1e3), "steps": args.codes_steps,
                          "output_bytes": P * S, "mode": "uint8 log-quantized mean codes"}
        log(f"uint8 codes: {cm:.2f} ms/step")
        del d_q
    if args.summary_steps > 0:
        ctx.run_fb_summary(pairs, core_block_sites=core, flank_sites=flank)
        t1 = time.perf_counter()
        for _ in range(args.summary_steps):
            ctx.run_fb_summary(pairs, core_block_sites=core, flank_sites=flank)
        sm = (time.perf_counter() - t1) * 1e3 / args.summary_steps
        ex["site_summary"] = {"ms": sm, "pair_sites_per_s": pair_sites / (sm / 1e3), "steps": args.summary_steps,
                              "timing": "host wall clock per call (pair upload, decode, 3 x n_sites floats back)"}
        log(f"summary mode: {sm:.2f} ms/call")
    return ex


def host_stream_run(ctx, pairs, core, flank, sample, alone, rows):
    """The PCIe-inclusive path infer_blockwise uses: decode into host arrays
    through device output slots and pinned staging buffers (device decode,
    D2H copies and the host scatter overlap). Returns its rate and checks
    the sampled columns against the device-resident run."""
    P, S = pairs.shape[0], alone.shape[0]
    t0 = time.perf_counter()
    r = ctx.run_fb_blockwise(pairs, core_block_sites=core, flank_sites=flank, pair_batch_size=-1, max_streams=2)
    dt = time.perf_counter() - t0
    same = bits_equal(r["mean"][np.asarray(rows)][:, sample], alone[np.asarray(rows)])
    del r
    gc.collect()
    gb = 4.0 * P * S / 1e9
    log(f"host stream: {dt:.2f}s for {gb:.1f} GB of output ({gb / dt:.1f} GB/s)")
    return {"seconds": dt, "pair_sites_per_s": float(P) * S / dt, "output_gb": gb, "output_gb_per_s": gb / dt,
            "max_streams": 2, "checks": {"sample_matches_device_run": same}}


def run_ragged(args, wl, compact=False):
    """Batched ragged form: every job's (pair tile x block window) items in
    one launch, against the plain mode (one launch per job, each waited
    for). Both write the same buffers, and the outputs must be identical."""
    from gamma_smc_cu_amd import _core, autotune, synthetic
    from gamma_smc_cu_amd.batch import Job, Staged, Workspace
    from gamma_smc_cu_amd.device import Event, Stream, synchronize
    from gamma_smc_cu_amd.flow_field import DEFAULT_PATH
    from gamma_smc_cu_amd.infer import _estimate_scaled_params

    n, S = wl["n"], wl["S"]
    t0 = time.perf_counter()
    G, pos = synthetic.cohort(n, S, wl["span"], wl["seed"])
    raw = synthetic.ragged_jobs(n, S, wl["n_jobs"], wl["seed"])
    log(f"ragged: cohort {n} x {S} and {len(raw)} jobs in {time.perf_counter() - t0:.1f}s")
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
        log(f"ragged: autotune picked {tuning}")
    else:
        tuning = dict(checkpoint_k=args.checkpoint_k, pairs_per_lane=args.pairs_per_lane, schedule=args.schedule,
                      xcd_order=bool(args.xcd_order))
        ctx.set_tuning(**tuning)
    steps = 3 if compact else args.steps
    for _ in range(1 if compact else max(1, args.warmup)):
        batched()
    synchronize()
    times = [batched() for _ in range(steps)]
    ms = sum(times) / len(times)
    log(f"ragged: batched {ms:.2f} ms/launch over {steps} launches")

    # Snapshot rows of every job from the batched run, then rerun plain
    # into the same buffers and compare bit for bit.
    def rows_of_job(k):
        j = jobs[k]
        P = int(st.pair_off[k + 1] - st.pair_off[k])
        out = []
        for r in sorted({0, j.n_sites // 2, j.n_sites - 1}):
            buf = np.empty(P, np.float32)
            _core.check(_core.lib().gsmc_memcpy_dtoh(ctypes.c_void_p(buf.ctypes.data),
                                                     st.d_out.offset(int(st.out_off[k]) + r * P), P * 4))
            out.append(buf)
        return np.concatenate(out)

    snap = [rows_of_job(k) for k in range(len(jobs))]
    plain_ev, plain_wall = plain()
    plain_ev2, plain_wall2 = plain()
    plain_ms = min(plain_ev, plain_ev2)
    log(f"ragged: plain {plain_ms:.2f} ms (events), wall {min(plain_wall, plain_wall2):.2f} ms for {len(jobs)} "
        f"launches")
    checks = {"batched_equals_plain_bit_exact": all(bits_equal(snap[k], rows_of_job(k)) for k in range(len(jobs)))}

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
        ref = O.fb_blockwise_f32(np.ascontiguousarray(G[:, j.site_lo:j.site_hi]), pos[j.site_lo:j.site_hi], sp, T, Ne,
                                 wl["core"], wl["flank"])
        cpu_s = time.perf_counter() - t1
        P = len(j.pairs)
        full = np.empty((j.n_sites, P), np.float32)
        _core.check(_core.lib().gsmc_memcpy_dtoh(ctypes.c_void_p(full.ctypes.data),
                                                 st.d_out.offset(int(st.out_off[k])), full.nbytes))
        checks["oracle_bit_exact_sampled_job"] = bits_equal(ref["mean"], full[:, :m])
        info = cpu_info()
        cpu = {"value": m * j.n_sites / cpu_s, "unit": "pair-sites/s", "cores": min(info["omp_threads"], m),
               "kind": "port", "sample": f"{m} pairs x {j.n_sites} sites of job {k} (its blocks), {cpu_s:.2f}s",
               **info}

    K = tuning["checkpoint_k"]
    B = sum(len(j.pairs) * j.n_sites * (4.0 + 16.0 / K) + 8 * j.n_sites + 8 * len(j.pairs) for j in jobs) \
        + n * ((S + 63) // 64) * 8
    achieved = B / (ms / 1e3) / 1e9
    out = {"label": wl["label"], "n_jobs": len(jobs), "pair_sites": pair_sites, "tuning": tuning, "ms": ms,
           "value": pair_sites / (ms / 1e3), "steps": steps,
           "plain": {"value": pair_sites / (plain_ms / 1e3), "ms": plain_ms, "launches": len(jobs)},
           "speedup_vs_plain": plain_ms / ms,
           "roofline": {"bound": "ta", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                        "kernel": "fb_fwd + fb_bwd (one batched launch pair)", "algorithmic_bytes_per_launch": B},
           "cpu_baseline": cpu, "checks": checks, "autotune": sweep,
           "sweep_note": "pairs per workgroup 64 / 128 / 256 = pairs_per_lane 1 / 2 / 4 (a workgroup is one "
                         "64-lane wave); checkpoint_k is the recomputed register window (1, 8, 16), not the "
                         "reference's K in {64, 256, 1024}"}
    del st, ws
    ctx.close()
    gc.collect()
    return out


def run_serving(args, wl, compact=False):
    """Serving loop: Poisson query streams at swept rates against the
    resident cohort, continuous batching vs one query per launch. Latency
    is arrival to device completion of the query's launch."""
    from gamma_smc_cu_amd import _core, serving, synthetic
    from gamma_smc_cu_amd.flow_field import DEFAULT_PATH
    from gamma_smc_cu_amd.infer import _estimate_scaled_params_packed

    n, S = wl["n"], wl["S"]
    t0 = time.perf_counter()
    packed, pos = synthetic.resident_cohort(n, S, wl["span"], wl["seed"])
    log(f"serving: resident cohort {n} x {S} ({packed.nbytes / 1e6:.0f} MB packed) in "
        f"{time.perf_counter() - t0:.1f}s")
    Ne = 1e4
    kmu, krho = _estimate_scaled_params_packed(packed, n, pos, 1.25e-8, 1e-8, Ne)
    _core.set_device(0)
    ctx = _core.FlowContext(packed, pos, Ne, kmu, krho, DEFAULT_PATH, 0, packed=True, n_sites=S)
    del packed
    # checkpoints every 8 sites, static grid (the chromosome tuner's usual
    # pick) unless the command line fixes a shape
    tuning = (dict(checkpoint_k=8, pairs_per_lane=1, schedule="static", xcd_order=False) if args.tune == "auto"
              else dict(checkpoint_k=args.checkpoint_k, pairs_per_lane=args.pairs_per_lane, schedule=args.schedule,
                        xcd_order=bool(args.xcd_order)))
    ctx.set_tuning(**tuning)
    arena = int(wl["arena_gb"] * 1e9)
    rates = (800.0, 3200.0) if compact else wl["rates"]
    seconds = 1.5 if compact else wl["seconds"]
    max_q = 400 if compact else wl["max_queries"]

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
        log(f"serving {mode:8s} rate {rate:7.1f}/s: p50 {r['p50_ms']:8.2f} ms p99 {r['p99_ms']:8.2f} ms "
            f"{r['posteriors_per_s']:.3e} pair-sites/s, {srv.launches} launches")
        del srv
        return r

    # warm both paths (first launches of a shape, page-in)
    one_run("batched", rates[0], 8)
    one_run("single", rates[0], 4)
    rows = []
    for rate in rates:
        nq = int(min(max_q, max(40, rate * seconds)))
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
        same &= bits_equal(alone.result(c), res[q.qid])
        alone.release(c)
    top = max(rates)
    b = next(r for r in rows if r["mode"] == "batched" and r["rate_per_s"] == top)
    s1 = next(r for r in rows if r["mode"] == "single" and r["rate_per_s"] == top)
    out = {"label": wl["label"], "tuning": tuning, "arena_gb": wl["arena_gb"], "ck_gb_per_launch": wl["ck_gb"],
           "slots": 2, "latency": "query arrival to device completion of its launch",
           "value": b["posteriors_per_s"], "p50_ms": b["p50_ms"], "p99_ms": b["p99_ms"], "top_rate_per_s": top,
           "single_query_per_launch": {"value": s1["posteriors_per_s"], "p50_ms": s1["p50_ms"],
                                       "p99_ms": s1["p99_ms"]},
           "speedup_vs_single": b["posteriors_per_s"] / s1["posteriors_per_s"], "sweep": rows,
           "checks": {"alone_vs_mix_bit_exact": same}}
    del mix, alone
    ctx.close()
    gc.collect()
    return out


def emit(line, checks):
    print(json.dumps(line), flush=True)
    if not all(checks.values()):
        sys.exit(1)


def all_checks(d):
    out = {}
    for k, v in (d.get("checks") or {}).items():
        out[k] = v
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="chromosome", choices=sorted(WORKLOADS))
    ap.add_argument("--tune", default="auto", choices=("auto", "none"),
                    help="auto: sweep launch shapes first and keep the fastest (results are identical)")
    ap.add_argument("--checkpoint-k", type=int, default=8)
    ap.add_argument("--pairs-per-lane", type=int, default=1)
    ap.add_argument("--schedule", default="static", choices=("static", "persistent", "dynamic"))
    ap.add_argument("--xcd-order", type=int, default=0)
    ap.add_argument("--cpu-pairs", type=int, default=1000)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-checks", action="store_true",
                    help="skip the sampled-pairs checks (profiling runs: only same-shape decode launches)")
    ap.add_argument("--no-subresults", action="store_true",
                    help="chromosome only: skip the end-to-end, biobank, ragged and serving sub-results")
    ap.add_argument("--tuning-json", default=None,
                    help="take the launch shape from the config.tuning of a previous bench JSON line")
    ap.add_argument("--graph-steps", type=int, default=2)
    ap.add_argument("--codes-steps", type=int, default=2)
    ap.add_argument("--fwd-steps", type=int, default=2)
    ap.add_argument("--summary-steps", type=int, default=2)
    ap.add_argument("--gene-steps", type=int, default=2)
    ap.add_argument("--no-host-stream", action="store_true")
    ap.add_argument("--copy-calibration-gib", type=float, default=1.0,
                    help="size of the plain copy kernel run after the timed region (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 "
                                   "(the decoder would shard over pairs with no communication)"}))
        sys.exit(2)

    wl = WORKLOADS[args.workload]
    base = {"n_gpus": 1, "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
            "data": "synthetic"}
    if args.workload == "ragged":
        r = run_ragged(args, wl)
        line = {"metric": "pair-site posteriors/sec (batched ragged jobs)", "value": r["value"],
                "unit": "pair-sites/s", "steps": r["steps"], "warmup": args.warmup, "ms_per_step": r["ms"], **base,
                "config": {"workload": r["label"], "tuning": r["tuning"], "parallelism": "single GPU"}, **r}
        emit(line, r["checks"])
        return
    if args.workload == "serving":
        r = run_serving(args, wl)
        line = {"metric": "serving pair-site posteriors/sec at the top offered rate (p50/p99 latency reported)",
                "value": r["value"], "unit": "pair-sites/s", "steps": len(r["sweep"]), "warmup": 2,
                "ms_per_step": None, **base, "config": {"workload": r["label"], "parallelism": "single GPU"},
                "roofline": None, "cpu_baseline": None, **r}
        emit(line, r["checks"])
        return

    head, state = run_decode_workload(args, args.workload)
    checks = dict(head["checks"])
    host = None
    if args.workload == "biobank" and state and not args.no_host_stream:
        host = host_stream_run(state["ctx"], state["pairs"], head["core"], head["flank"], state["sample"],
                               state["alone"], state["rows"])
        checks.update({f"host_stream_{k}": v for k, v in host["checks"].items()})
    sub = {}
    if args.workload == "chromosome" and not args.no_subresults:
        if state:
            G, pos, sample, alone, rows = state["G"], state["pos"], state["sample"], state["alone"], state["rows"]
            pairs = state["pairs"]
            state["ctx"].close()
            state = None
            gc.collect()
            sub["end_to_end"] = end_to_end(G, pos, pairs[sample], alone, rows)
            del G, pos, alone
            gc.collect()
        b, bstate = run_decode_workload(args, "biobank", compact=True)
        if bstate:
            b["host_stream"] = host_stream_run(bstate["ctx"], bstate["pairs"], b["core"], b["flank"],
                                               bstate["sample"], bstate["alone"], bstate["rows"])
            b["checks"].update({f"host_stream_{k}": v for k, v in b["host_stream"]["checks"].items()})
            bstate["ctx"].close()
            bstate = None
            gc.collect()
        sub["biobank"] = b
        sub["ragged"] = run_ragged(args, WORKLOADS["ragged"], compact=True)
        sub["serving"] = run_serving(args, WORKLOADS["serving"], compact=True)
        sub["plain"] = plain_small()
        sub["scaling"] = scaling_infer()
        for name, r in sub.items():
            checks.update({f"{name}.{k}": v for k, v in (r.get("checks") or {}).items()})

    line = {
        "metric": "pair-site posteriors/sec",
        "value": head["value"],
        "unit": "pair-sites/s",
        "steps": head["steps"],
        "warmup": head["warmup"],
        "ms_per_step": head["ms"],
        **base,
        "config": {"workload": head["label"], "n_haplotypes": head["n"], "n_pairs": head["P"],
                   "n_sites": head["S"], "core_block_sites": head["core"], "flank_sites": head["flank"],
                   "n_blocks": head["n_blocks"], "tuning": head["tuning"], "mode": "mean_only",
                   "parallelism": "single GPU"},
        "roofline": head["roofline"],
        "cpu_baseline": head.get("cpu_baseline"),
        "fp64_error": head.get("fp64_error"),
        "host_stream": host,
        "site_summary": head.get("site_summary"),
        "gene_stats": head.get("gene_stats"),
        "forward_only": head.get("forward_only"),
        "codes_u8": head.get("codes_u8"),
        "graph_replay": head.get("graph_replay"),
        "checks": checks,
        "wall_ms_per_step": head["wall_ms_per_step"],
        "copy_calibration_bytes": head["copy_bytes"],
        "autotune": head["autotune"],
        **sub,
    }
    emit(line, checks)


if __name__ == "__main__":
    main()
