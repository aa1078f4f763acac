"""GPU tests at the benchmark workloads' own shapes.

1. Plain ``infer()`` on a 20-haplotype, 5 Mb cohort (all 190 pairs, ~9k
   sites, integer positions): every output against the FP32 oracle, in
   full. msprime is not installed, so the cohort is the synthetic 1/k one
   at the expected site count (4 Ne mu L H_19 ~ 8.9k); parity against an
   msprime-generated input is unpinned.
2. The chromosome: 356 haplotypes (all 63,190 pairs) x 402,853 sites,
   decoded on the device at full size with the auto-sized core blocks and
   2048-site flanks; 64 random pairs decoded alone equal the oracle over
   every site, and equal the full run's columns at sampled rows.
3. 64 ragged jobs over a 2,000-haplotype cohort in one batched launch equal
   the plain one-launch-per-job mode, and the oracle on sampled pairs.
4. The serving loop over the resident 6,404 x 1,000,000 cohort: a Poisson
   burst, each result equal alone and in the mix, and the oracle on a few
   queries.
5. The biobank cohort: 1,000 haplotypes (499,500 pairs) x 50,000 sites at
   full size; a 1,000-pair sample within the stated FP64 error bounds and
   bit-exact to the FP32 oracle; the host-output path (device slots,
   pinned staging) equal to the device run.
"""

import ctypes
import gc
import importlib

import numpy as np
import pytest

import gamma_smc_cu_amd
from gamma_smc_cu_amd import _core, serving, synthetic
from gamma_smc_cu_amd.batch import Job, Staged, Workspace
from gamma_smc_cu_amd.device import DeviceArray, Stream, run_fb_device
from oracle import oracle as O

pytestmark = pytest.mark.gpu
infer_mod = importlib.import_module("gamma_smc_cu_amd.infer")
NE = 1e4


def _case(msg):
    print(f"[case] {msg}", flush=True)


def _bits_equal(a, b):
    a, b = np.ascontiguousarray(a), np.ascontiguousarray(b)
    return a.shape == b.shape and np.array_equal(a.view(np.uint32), b.view(np.uint32))


def _rows(d_out, rows, P, cols=None):
    """Rows of a device (S, P) float32 array (optionally the first `cols` columns)."""
    n = P if cols is None else cols
    out = np.empty((len(rows), n), np.float32)
    for k, r in enumerate(rows):
        _core.check(_core.lib().gsmc_memcpy_dtoh(ctypes.c_void_p(out[k].ctypes.data), d_out.offset(int(r) * P), n * 4))
    return out


def _all_pairs_array(n):
    a, b = np.tril_indices(n, k=-1)
    return np.stack([a, b], axis=1).astype(np.int32)


def _decode_full_and_sample(ctx, G, pos, P, S, core, flank, sample, flow_field_path, mu, rho):
    """Decode every pair on the device, then the sampled pairs alone; return
    (alone (S, m), sampled rows of the full run (rows, m), the rows)."""
    pairs = _all_pairs_array(G.shape[0])
    assert pairs.shape[0] == P
    st = Stream()
    d_pi = DeviceArray.from_host(np.ascontiguousarray(pairs[:, 0]))
    d_pj = DeviceArray.from_host(np.ascontiguousarray(pairs[:, 1]))
    d_out = DeviceArray((S, P), np.float32)
    run_fb_device(ctx, d_pi, d_pj, P, core, flank, 0, d_out, stream=st)
    sp = pairs[sample]
    d_spi = DeviceArray.from_host(np.ascontiguousarray(sp[:, 0]))
    d_spj = DeviceArray.from_host(np.ascontiguousarray(sp[:, 1]))
    d_sout = DeviceArray((S, len(sample)), np.float32)
    run_fb_device(ctx, d_spi, d_spj, len(sample), core, flank, 0, d_sout, stream=st)
    st.synchronize()
    alone = d_sout.to_host()
    rng = np.random.default_rng(5)
    rows = np.unique(np.concatenate([[0, core - 1, min(core, S - 1), S // 2, S - 1], rng.integers(0, S, 24)]))
    full_rows = _rows(d_out, rows, P)
    d_out.free()
    return alone, full_rows, rows, pairs


def test_config1_plain_infer_all_190_pairs(flow_field_path):
    G, pos = synthetic.cohort(20, 8_900, 5_000_000, 1)
    pos_i64 = pos.astype(np.int64)
    pairs = O.all_pairs(20)
    mu, rho = O.estimate_scaled_params(G, pos, 1.25e-8, 1e-8, NE)
    T = O.tables_f32(flow_field_path, 1000, NE, mu, rho)
    _case("config 1: infer, 190 pairs x 8,900 sites, mean only")
    r = gamma_smc_cu_amd.infer(G, pos_i64, mu=1.25e-8, rho=1e-8)
    assert r["pairs"] == pairs and r["mean"].shape == (8_900, 190)
    ref = O.fb_window_f32(G, pos, pairs, T, NE, ci=True, posterior=True)
    assert _bits_equal(r["mean"], ref["mean"])
    _case("config 1: infer with CI and posterior")
    r = gamma_smc_cu_amd.infer(G, pos_i64, mu=1.25e-8, rho=1e-8, mean_only=False, return_posterior=True)
    for k in ("mean", "lower", "upper", "posterior_alpha", "posterior_beta"):
        assert _bits_equal(r[k], ref[k]), k


def test_config2_chromosome_full_size(flow_field_path):
    n, S = 356, 402_853
    _case("config 2: cohort 356 x 402,853")
    G, pos = synthetic.cohort(n, S, 51_000_000, 7)
    P = n * (n - 1) // 2
    mu, rho = infer_mod._estimate_scaled_params(G, pos, 1.25e-8, 1e-8, NE)
    free, _ = _core.cuda_mem_info()
    core, _ = infer_mod._recommend_core_block_size(S, P, True, 2048, free)
    ctx = gamma_smc_cu_amd.FlowContext(G, pos, NE, mu, rho, flow_field_path, 0)
    sample = np.sort(np.random.default_rng(7).choice(P, size=64, replace=False))
    _case(f"config 2: full decode, {P} pairs, core {core}, flank 2048")
    alone, full_rows, rows, pairs = _decode_full_and_sample(ctx, G, pos, P, S, core, 2048, sample, flow_field_path,
                                                            mu, rho)
    assert _bits_equal(full_rows[:, sample], alone[rows])
    assert np.isfinite(full_rows).all() and (full_rows > 0).all()
    _case("config 2: oracle on 64 pairs x all sites")
    T = O.tables_f32(flow_field_path, 1000, NE, mu, rho)
    ref = O.fb_blockwise_f32(G, pos, [tuple(map(int, p)) for p in pairs[sample]], T, NE, core, 2048)
    assert _bits_equal(alone, ref["mean"])
    ctx.close()
    gc.collect()


def test_config3_ragged_batched_equals_plain_and_oracle(flow_field_path):
    n, S = 2000, 262_144
    _case("config 3: cohort 2,000 x 262,144 and 64 jobs")
    G, pos = synthetic.cohort(n, S, S * 150, 11)
    raw = synthetic.ragged_jobs(n, S, 64, 11)
    mu, rho = infer_mod._estimate_scaled_params(G, pos, 1.25e-8, 1e-8, NE)
    ctx = gamma_smc_cu_amd.FlowContext(G, pos, NE, mu, rho, flow_field_path, 0)
    ctx.set_tuning(checkpoint_k=8)
    jobs = [Job(lo, hi, p, core_block_sites=32_768, flank_sites=2048) for lo, hi, p in raw]
    st = Staged(jobs)
    ws = Workspace(ctx)
    arr, nj = st.structs()
    _case("config 3: one batched launch")
    ws.launch(arr, nj, st.flags)
    ws.wait()

    def rows_of(k):
        j = jobs[k]
        P = len(j.pairs)
        rr = sorted({0, j.n_sites // 3, j.n_sites // 2, j.n_sites - 1})
        out = np.empty((len(rr), P), np.float32)
        for i, r in enumerate(rr):
            _core.check(_core.lib().gsmc_memcpy_dtoh(ctypes.c_void_p(out[i].ctypes.data),
                                                     st.d_out.offset(int(st.out_off[k]) + r * P), P * 4))
        return out

    batched = [rows_of(k) for k in range(len(jobs))]
    _case("config 3: plain mode, one launch per job")
    for k in range(len(jobs)):
        a, m = st.structs([k])
        ws.launch(a, m, st.flags)
        ws.wait()
    for k in range(len(jobs)):
        assert _bits_equal(rows_of(k), batched[k]), k
    _case("config 3: oracle on sampled pairs of two jobs")
    T = O.tables_f32(flow_field_path, 1000, NE, mu, rho)
    small = [k for k in range(len(jobs)) if len(jobs[k].pairs) <= 1000]
    for k in (max(small, key=lambda i: jobs[i].n_sites), min(small, key=lambda i: jobs[i].n_sites)):
        j = jobs[k]
        P = len(j.pairs)
        full = np.empty((j.n_sites, P), np.float32)
        _core.check(_core.lib().gsmc_memcpy_dtoh(ctypes.c_void_p(full.ctypes.data),
                                                 st.d_out.offset(int(st.out_off[k])), full.nbytes))
        m = min(96, P)
        sp = [tuple(map(int, p)) for p in j.pair_array()[:m]]
        ref = O.fb_blockwise_f32(np.ascontiguousarray(G[:, j.site_lo:j.site_hi]), pos[j.site_lo:j.site_hi], sp, T,
                                 NE, 32_768, 2048)
        assert _bits_equal(full[:, :m], ref["mean"]), k
    del st
    ctx.close()
    gc.collect()


def _unpack_rows(packed, haps, lo, hi):
    w0, w1 = lo // 64, (hi + 63) // 64
    words = np.ascontiguousarray(packed[np.asarray(haps), w0:w1])
    bits = np.unpackbits(words.view(np.uint8), axis=1, bitorder="little")
    return np.ascontiguousarray(bits[:, lo - w0 * 64:hi - w0 * 64]).astype(np.uint8)


def test_config4_serving_resident_cohort(flow_field_path):
    n, S = 6404, 1_000_000
    _case("config 4: resident cohort 6,404 x 1,000,000")
    packed, pos = synthetic.resident_cohort(n, S, 150_000_000, 13)
    mu, rho = infer_mod._estimate_scaled_params_packed(packed, n, pos, 1.25e-8, 1e-8, NE)
    ctx = gamma_smc_cu_amd.FlowContext(packed, pos, NE, mu, rho, flow_field_path, 0, packed=True, n_sites=S)
    qs, off = serving.poisson_queries(24, 400.0, n, S, 13, subset=(4, 60), core_block_sites=8192)
    srv = serving.Server(ctx, 16 << 30, n_slots=2, ck_bytes_per_launch=4 << 30)
    got = {}

    def keep(q):
        got[q.qid] = srv.result(q)
        srv.release(q)

    _case("config 4: Poisson burst, 24 queries at 400/s, 2 slots")
    serving.run_open_loop(srv, qs, off, keep)
    assert len(got) == len(qs) and srv.arena.used == 0
    lat = np.array([q.latency for q in qs])
    assert np.isfinite(lat).all() and (lat > 0).all()
    _case("config 4: each query alone")
    alone = serving.Server(ctx, 8 << 30, n_slots=1, ck_bytes_per_launch=4 << 30, max_queries_per_launch=1)
    for q in qs:
        c = serving.Query(q.haplotypes, q.site_lo, q.site_hi, q.core_block_sites, q.flank_sites)
        alone.submit(c)
        alone.drain()
        assert _bits_equal(alone.result(c), got[q.qid]), q.qid
        alone.release(c)
    _case("config 4: oracle on three queries")
    T = O.tables_f32(flow_field_path, 1000, NE, mu, rho)
    for q in sorted(qs, key=lambda q: q.pair_sites)[:3]:
        haps = np.unique(np.asarray(q.haplotypes))
        Gq = _unpack_rows(packed, haps, q.site_lo, q.site_hi)
        local = {int(h): k for k, h in enumerate(haps)}
        pairs = [(local[int(a)], local[int(b)]) for a, b in q.pairs]
        ref = O.fb_blockwise_f32(Gq, pos[q.site_lo:q.site_hi], pairs, T, NE, q.core_block_sites, q.flank_sites)
        assert _bits_equal(got[q.qid], ref["mean"]), q.qid
    del srv, alone
    ctx.close()
    gc.collect()


def test_config5_biobank_full_size(flow_field_path):
    n, S = 1000, 50_000
    _case("config 5: cohort 1,000 x 50,000")
    G, pos = synthetic.cohort(n, S, 7_500_000, 17)
    P = n * (n - 1) // 2
    mu, rho = infer_mod._estimate_scaled_params(G, pos, 1.25e-8, 1e-8, NE)
    free, _ = _core.cuda_mem_info()
    core, _ = infer_mod._recommend_core_block_size(S, P, True, 2048, free)
    ctx = gamma_smc_cu_amd.FlowContext(G, pos, NE, mu, rho, flow_field_path, 0)
    sample = np.sort(np.random.default_rng(17).choice(P, size=1000, replace=False))
    _case(f"config 5: full decode, {P} pairs, core {core}")
    alone, full_rows, rows, pairs = _decode_full_and_sample(ctx, G, pos, P, S, core, 2048, sample, flow_field_path,
                                                            mu, rho)
    assert _bits_equal(full_rows[:, sample], alone[rows])
    assert np.isfinite(full_rows).all() and (full_rows > 0).all()
    sp = [tuple(map(int, p)) for p in pairs[sample]]
    _case("config 5: FP64 oracle on 1,000 pairs")
    T64 = O.tables_f64(flow_field_path, 1000, NE, mu, rho)
    ref64 = O.fb_blockwise_f64(G, pos, sp, T64, NE, core, 2048)["mean"]
    err = np.abs(np.log(alone.astype(np.float64)) - np.log(ref64))
    assert err.max() <= 1e-2 and np.median(err) <= 1e-3, (err.max(), np.median(err))
    _case("config 5: FP32 oracle on 128 pairs")
    T = O.tables_f32(flow_field_path, 1000, NE, mu, rho)
    ref = O.fb_blockwise_f32(G, pos, sp[:128], T, NE, core, 2048)["mean"]
    assert _bits_equal(alone[:, :128], ref)
    _case("config 5: host-output path, 16,384 pairs, two slots")
    host = ctx.run_fb_blockwise(pairs[:16_384], core_block_sites=core, flank_sites=2048, pair_batch_size=-1,
                                max_streams=2)["mean"]
    st = Stream()
    d_pi = DeviceArray.from_host(np.ascontiguousarray(pairs[:16_384, 0]))
    d_pj = DeviceArray.from_host(np.ascontiguousarray(pairs[:16_384, 1]))
    d_out = DeviceArray((S, 16_384), np.float32)
    run_fb_device(ctx, d_pi, d_pj, 16_384, core, 2048, 0, d_out, stream=st)
    st.synchronize()
    assert _bits_equal(host, d_out.to_host())
    ctx.close()
    gc.collect()
