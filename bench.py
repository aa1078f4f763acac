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
  OpenMP over pairs) timed on this host on a 1000-pair sample over all
  sites, extrapolated per pair-site;
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

WORKLOADS = {
    "chromosome": dict(n=356, S=402_853, span=51_000_000, seed=7, flank=2048,
                       label="356 haplotypes (63,190 pairs) x 402,853 sites, 1/k spectrum, "
                             "infer_blockwise mean_only, auto core + 2048-site flanks"),
    "biobank": dict(n=1000, S=50_000, span=7_500_000, seed=17, flank=2048,
                    label="1,000 haplotypes (499,500 pairs) x 50,000 sites, 1/k spectrum, "
                          "infer_blockwise mean_only, auto core + 2048-site flanks"),
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

    return checks, cpu


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
    ap.add_argument("--copy-calibration-gib", type=float, default=1.0,
                    help="size of the plain copy kernel run after the timed region (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 "
                                   "(the decoder shards over pairs with no communication; see DESIGN.md)"}))
        sys.exit(2)

    from gamma_smc_cu_amd import _core, synthetic
    from gamma_smc_cu_amd.device import DeviceArray, Event, Stream, run_fb_device, synchronize
    from gamma_smc_cu_amd.flow_field import DEFAULT_PATH
    from gamma_smc_cu_amd.infer import _estimate_scaled_params, _recommend_core_block_size

    wl = WORKLOADS[args.workload]
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

    checks = {}
    cpu = None
    if not args.no_checks:
        checks, cpu = sample_checks(args, ctx, G, pos, pairs, P, S, core, flank, Ne, kmu, krho, d_out, stream)

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
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": ("fb_fwd_items + fb_bwd_items" if tuning["schedule"] == "static"
                                else "fb_fwd_queue + fb_bwd_queue"),
                     "algorithmic_bytes_per_launch": B},
        "cpu_baseline": cpu,
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
