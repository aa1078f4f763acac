"""GPU parity: the HIP path against the CPU oracle, through the C ABI.

The oracle (``oracle/gsmc_oracle.c``) restates the reference's recursion in
FP32 operation for operation, so the comparison is bit-exact
(``np.array_equal`` on the raw float32 bits) for mean, CI and posterior
arrays, for full-sequence and blockwise decoding.
"""

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _case(msg):
    print(f"[case] {msg}", flush=True)


def _bits_equal(a, b):
    a = np.ascontiguousarray(a, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    return a.shape == b.shape and np.array_equal(a.view(np.uint32), b.view(np.uint32))


@pytest.fixture(scope="module")
def core():
    from gamma_smc_cu_amd import _core

    assert _core.device_count() >= 1
    return _core


def _tables(path, mu, rho, Ne=1e4, N=1000):
    return O.tables_f32(path, N, Ne, mu, rho)


def test_bitpack_roundtrip_matches_closed_form(core):
    rng = np.random.default_rng(0)
    for n, S in [(1, 1), (4, 1), (10, 64), (5, 128), (10, 100), (7, 1000), (3, 4097)]:
        _case(f"bitpack n={n} S={S}")
        G = rng.integers(0, 2, size=(n, S), dtype=np.uint8)
        pk = core.bitpack(G)
        assert pk.shape == (n, (S + 63) // 64)
        np.testing.assert_array_equal(pk, O.bitpack(G))
        np.testing.assert_array_equal(core.unpack(pk, n, S), G)
    _case("bitpack uses bit 0 of each byte")
    G = np.array([[2, 3, 255, 0]], np.uint8)
    np.testing.assert_array_equal(core.unpack(core.bitpack(G), 1, 4), G & 1)


@pytest.mark.parametrize("mean_only,posterior", [(True, False), (False, False), (True, True), (False, True)])
def test_run_fb_bit_exact(core, cohort_small, flow_field_path, mean_only, posterior):
    G, pos = cohort_small
    mu, rho = O.estimate_scaled_params(G, pos, 1.25e-8, 1e-8, 1e4)
    T = _tables(flow_field_path, mu, rho)
    pairs = O.all_pairs(G.shape[0])
    _case(f"run_fb mean_only={mean_only} posterior={posterior} P={len(pairs)} S={G.shape[1]}")
    ctx = core.FlowContext(G, pos, 1e4, mu, rho, flow_field_path, 0)
    got = ctx.run_fb(pairs, mean_only=mean_only, return_posterior=posterior)
    ref = O.fb_window_f32(G, pos, pairs, T, 1e4, ci=not mean_only, posterior=posterior)
    for k, v in ref.items():
        assert _bits_equal(got[k], v), k


@pytest.mark.parametrize("K", [8, 16])
def test_run_fb_checkpoint_interval_invariant(core, cohort_small, flow_field_path, K):
    G, pos = cohort_small
    pairs = [(1, 0), (5, 2), (11, 3), (4, 4)]
    _case(f"run_fb K={K}")
    ctx = core.FlowContext(G, pos, 1e4, 1.25e-8, 1e-8, flow_field_path, 0)
    ctx.set_tuning(checkpoint_k=K, min_waves_per_launch=1)
    got = ctx.run_fb(pairs, mean_only=False, return_posterior=True)
    ref = O.fb_window_f32(G, pos, pairs, _tables(flow_field_path, 1.25e-8, 1e-8), 1e4, ci=True, posterior=True)
    for k, v in ref.items():
        assert _bits_equal(got[k], v), k


@pytest.mark.parametrize("core_sites,flank", [(256, 64), (100, 37), (700, 0), (64, 300), (1, 5)])
def test_blockwise_bit_exact(core, cohort_small, flow_field_path, core_sites, flank):
    G, pos = cohort_small
    pairs = [(i, j) for i in range(G.shape[0]) for j in range(i)][:50]
    _case(f"blockwise core={core_sites} flank={flank}")
    ctx = core.FlowContext(G, pos, 1e4, 1.25e-8, 1e-8, flow_field_path, 0)
    got = ctx.run_fb_blockwise(pairs, core_block_sites=core_sites, flank_sites=flank, pair_batch_size=-1,
                               max_streams=1, mean_only=False, return_posterior=True)
    ref = O.fb_blockwise_f32(G, pos, pairs, _tables(flow_field_path, 1.25e-8, 1e-8), 1e4, core_sites, flank,
                             ci=True, posterior=True)
    np.testing.assert_array_equal(got["blocks"], ref["blocks"])
    for k in ("mean", "lower", "upper", "posterior_alpha", "posterior_beta"):
        assert _bits_equal(got[k], ref[k]), k


def test_blockwise_pair_chunks_and_streams_invariant(core, cohort_small, flow_field_path):
    G, pos = cohort_small
    pairs = O.all_pairs(G.shape[0])
    ctx = core.FlowContext(G, pos, 1e4, 1.25e-8, 1e-8, flow_field_path, 0)
    _case("blockwise pair_batch_size=7 max_streams=3")
    a = ctx.run_fb_blockwise(pairs, core_block_sites=128, flank_sites=32, pair_batch_size=7, max_streams=3)
    _case("blockwise single launch")
    b = ctx.run_fb_blockwise(pairs, core_block_sites=128, flank_sites=32, pair_batch_size=-1, max_streams=1)
    assert _bits_equal(a["mean"], b["mean"])


@pytest.mark.parametrize("n_pairs", [128, 130, 129, 64])
def test_row_segment_stores_bit_exact(core, cohort_small, flow_field_path, n_pairs):
    """Full 64-pair slots store their mean rows through an LDS tile as 16-byte
    (P % 4 == 0) or 8-byte (P % 2 == 0) row segments, other slots lane by
    lane; every pair count lands on the oracle's values bit for bit."""
    G, pos = cohort_small
    base = O.all_pairs(G.shape[0])
    pairs = (base * 3)[:n_pairs]
    T = _tables(flow_field_path, 1.25e-8, 1e-8)
    ctx = core.FlowContext(G, pos, 1e4, 1.25e-8, 1e-8, flow_field_path, 0)
    for K in (8, 16):
        ctx.set_tuning(checkpoint_k=K)
        _case(f"row stores P={n_pairs} K={K}")
        got = ctx.run_fb_blockwise(pairs, core_block_sites=200, flank_sites=64, pair_batch_size=-1, max_streams=1)
        assert _bits_equal(got["mean"], O.fb_blockwise_f32(G, pos, pairs, T, 1e4, 200, 64)["mean"]), K


def test_gap_longer_than_cache(core, flow_field_path):
    # gaps beyond n_max_steps exercise the stretch-layer chunking
    rng = np.random.default_rng(5)
    G = rng.integers(0, 2, size=(6, 40), dtype=np.uint8)
    pos = np.cumsum(rng.choice([1, 2, 999, 1000, 1001, 2500, 7, 0], size=40)).astype(np.float64) + 1500.0
    pairs = O.all_pairs(6)
    _case("long gaps, cache_steps=1000")
    ctx = core.FlowContext(G, pos, 1e4, 1.25e-8, 1e-8, flow_field_path, 0)
    got = ctx.run_fb(pairs, mean_only=False, return_posterior=True)
    ref = O.fb_window_f32(G, pos, pairs, _tables(flow_field_path, 1.25e-8, 1e-8), 1e4, ci=True, posterior=True)
    for k, v in ref.items():
        assert _bits_equal(got[k], v), k
    _case("long gaps, cache_steps=37")
    ctx = core.FlowContext(G, pos, 1e4, 1.25e-8, 1e-8, flow_field_path, 37)
    got = ctx.run_fb_blockwise(pairs, core_block_sites=16, flank_sites=5, mean_only=True)
    ref = O.fb_blockwise_f32(G, pos, pairs, _tables(flow_field_path, 1.25e-8, 1e-8, N=37), 1e4, 16, 5)
    assert _bits_equal(got["mean"], ref["mean"])


def test_single_site_and_empty(core, flow_field_path):
    G = np.array([[0], [1], [1], [0]], np.uint8)
    pos = np.array([12345.0])
    ctx = core.FlowContext(G, pos, 1e4, 1.25e-8, 1e-8, flow_field_path, 0)
    _case("S=1")
    got = ctx.run_fb([(1, 0), (2, 1)], mean_only=False)
    ref = O.fb_window_f32(G, pos, [(1, 0), (2, 1)], _tables(flow_field_path, 1.25e-8, 1e-8), 1e4, ci=True)
    for k, v in ref.items():
        assert _bits_equal(got[k], v), k
    _case("P=0")
    assert ctx.run_fb([])["mean"].shape == (1, 0)
    r = ctx.run_fb_blockwise([], core_block_sites=8, flank_sites=2)
    assert r["mean"].shape == (1, 0) and r["blocks"].shape == (1, 4)


def test_pair_symmetry_self_pairs_and_two_haplotypes(core, cohort_small, flow_field_path):
    """A pair and its swap decode to the same bits (het calls are the XOR of
    the two rows); a self pair has no het site and equals the oracle; a
    two-haplotype cohort has its single pair. The reference's property
    suite checks the same symmetry and edge cases on its HMM path
    (tests/property/test_invariants.py:17-34, test_edge_cases.py:68-118)."""
    G, pos = cohort_small
    ctx = core.FlowContext(G, pos, 1e4, 1.25e-8, 1e-8, flow_field_path, 0)
    pairs = [(0, 1), (1, 0), (3, 3), (2, 7), (7, 2), (11, 0), (0, 11)]
    _case("pair symmetry")
    got = ctx.run_fb_blockwise(pairs, core_block_sites=200, flank_sites=64)["mean"]
    for a, b in ((0, 1), (3, 4), (5, 6)):
        assert _bits_equal(got[:, a], got[:, b]), (pairs[a], pairs[b])
    ref = O.fb_blockwise_f32(G, pos, [(3, 3)], _tables(flow_field_path, 1.25e-8, 1e-8), 1e4, 200, 64)["mean"]
    assert _bits_equal(got[:, 2:3], ref)
    _case("two haplotypes")
    G2 = np.ascontiguousarray(G[[4, 9]])
    keep = G2.any(axis=0) & ~G2.all(axis=0)
    G2, pos2 = np.ascontiguousarray(G2[:, keep]), pos[keep]
    ctx2 = core.FlowContext(G2, pos2, 1e4, 1.25e-8, 1e-8, flow_field_path, 0)
    got2 = ctx2.run_fb([(1, 0)], mean_only=False)
    ref2 = O.fb_window_f32(G2, pos2, [(1, 0)], _tables(flow_field_path, 1.25e-8, 1e-8), 1e4, ci=True)
    for k, v in ref2.items():
        assert _bits_equal(got2[k], v), k


def test_pair_index_out_of_range_raises(core, cohort_small, flow_field_path):
    G, pos = cohort_small
    ctx = core.FlowContext(G, pos, 1e4, 1.25e-8, 1e-8, flow_field_path, 0)
    with pytest.raises(ValueError, match="pair index out of range"):
        ctx.run_fb([(0, G.shape[0])])


SCHEDULES = [dict(checkpoint_k=k, pairs_per_lane=np_, schedule=sch, xcd_order=xcd)
             for k in (8, 16) for np_ in (1, 2)
             for sch, xcd in (("static", True), ("static", False), ("persistent", False))]
SCHEDULES += [dict(checkpoint_k=1, pairs_per_lane=1, schedule=sch, xcd_order=xcd)
              for sch, xcd in (("static", True), ("persistent", False))]
SCHEDULES += [dict(checkpoint_k=k, pairs_per_lane=1, schedule="dynamic", xcd_order=False) for k in (8, 16, 1)]


def test_launch_shapes_are_bit_identical(core, cohort_small, flow_field_path):
    G, pos = cohort_small
    pairs = O.all_pairs(G.shape[0]) + [(3, 3)]
    ref = O.fb_blockwise_f32(G, pos, pairs, _tables(flow_field_path, 1.25e-8, 1e-8), 1e4, 200, 64, ci=True,
                             posterior=True)
    ctx = core.FlowContext(G, pos, 1e4, 1.25e-8, 1e-8, flow_field_path, 0)
    for tu in SCHEDULES:
        _case(f"tuning {tu}")
        ctx.set_tuning(**tu, min_waves_per_launch=3)
        got = ctx.run_fb_blockwise(pairs, core_block_sites=200, flank_sites=64, pair_batch_size=-1,
                                   mean_only=False, return_posterior=True)
        for k in ("mean", "lower", "upper", "posterior_alpha", "posterior_beta"):
            assert _bits_equal(got[k], ref[k]), (tu, k)


def _ragged_jobs(n_haps, S, rng):
    from gamma_smc_cu_amd.batch import Job

    jobs = []
    for k in range(9):
        lo = int(rng.integers(0, S - 40))
        hi = int(rng.integers(lo + 1, min(S, lo + 400) + 1))
        npairs = int(rng.integers(1, 90))
        pairs = rng.integers(0, n_haps, size=(npairs, 2))
        core_sites = int(rng.choice([0, 50, 128]))
        flank = 0 if core_sites == 0 else int(rng.integers(0, 40))
        jobs.append(Job(lo, hi, pairs, core_sites, flank))
    jobs.append(Job(5, 5, [(0, 1)]))          # empty range
    jobs.append(Job(0, S, np.zeros((0, 2))))  # no pairs
    return jobs


@pytest.mark.parametrize("tu", [SCHEDULES[0], SCHEDULES[4], SCHEDULES[8], SCHEDULES[13], SCHEDULES[14],
                                SCHEDULES[16]])
def test_batch_matches_alone_and_oracle(core, cohort_small, flow_field_path, tu):
    from gamma_smc_cu_amd import batch

    G, pos = cohort_small
    rng = np.random.default_rng(17)
    jobs = _ragged_jobs(G.shape[0], G.shape[1], rng)
    ctx = core.FlowContext(G, pos, 1e4, 1.25e-8, 1e-8, flow_field_path, 0)
    ctx.set_tuning(**tu)
    _case(f"batch {tu}")
    mixed = batch.decode_batch(ctx, jobs, mean_only=False, return_posterior=True)
    _case("plain")
    alone = batch.decode_plain(ctx, jobs, mean_only=False, return_posterior=True)
    T = _tables(flow_field_path, 1.25e-8, 1e-8)
    for j, a, b in zip(jobs, mixed, alone):
        for k in a:
            assert _bits_equal(a[k], b[k]), k
        pairs = [tuple(map(int, p)) for p in j.pair_array()]
        if j.n_sites == 0 or not pairs:
            assert a["mean"].size == 0
            continue
        sub_G = np.ascontiguousarray(G[:, j.site_lo:j.site_hi])
        sub_pos = np.ascontiguousarray(pos[j.site_lo:j.site_hi])
        core_sites = j.core_block_sites if j.core_block_sites > 0 else j.n_sites
        flank = j.flank_sites if j.core_block_sites > 0 else 0
        ref = O.fb_blockwise_f32(sub_G, sub_pos, pairs, T, 1e4, core_sites, flank, ci=True, posterior=True)
        for k in ("mean", "lower", "upper", "posterior_alpha", "posterior_beta"):
            assert _bits_equal(a[k], ref[k]), k


@pytest.mark.parametrize("schedule,budget,np_", [("static", 0, 1), ("persistent", 0, 1), ("static", 400_000, 1),
                                                 ("persistent", 800_000, 1), ("static", 0, 2),
                                                 ("static", 800_000, 2), ("dynamic", 0, 1),
                                                 ("dynamic", 800_000, 2)])
def test_stored_forward_states_bit_exact(core, cohort_small, flow_field_path, schedule, budget, np_):
    """checkpoint_k=1: each wave's forward walk stores the state after
    every core site in the wave's slot and its backward walk reads it back
    instead of recomputing it. Means, log codes and ragged batches equal the
    oracle's bit for bit, also when a state budget of about one slot of the
    largest work item (700 sites x 64 pairs x 8 B = 358,400 B) leaves a
    single wave, or two, to decode every item in turn."""
    from gamma_smc_cu_amd import batch

    G, pos = cohort_small
    pairs = O.all_pairs(G.shape[0]) + [(3, 3)]
    T = _tables(flow_field_path, 1.25e-8, 1e-8)
    ctx = core.FlowContext(G, pos, 1e4, 1.25e-8, 1e-8, flow_field_path, 0)
    ctx.set_tuning(checkpoint_k=1, pairs_per_lane=np_, schedule=schedule)
    ctx.set_state_budget(budget)
    _case(f"stored states {schedule} budget={budget} pairs/lane={np_}: full sequence")
    got = ctx.run_fb(pairs, mean_only=True)
    assert _bits_equal(got["mean"], O.fb_window_f32(G, pos, pairs, T, 1e4)["mean"])
    for csz, fl in ((200, 64), (37, 5), (1, 3)):
        _case(f"stored states {schedule} budget={budget}: core={csz} flank={fl}")
        got = ctx.run_fb_blockwise(pairs, core_block_sites=csz, flank_sites=fl, pair_batch_size=-1, max_streams=1)
        ref = O.fb_blockwise_f32(G, pos, pairs, T, 1e4, csz, fl)["mean"]
        assert _bits_equal(got["mean"], ref), (csz, fl)
    _case(f"stored states {schedule} budget={budget}: u8 codes")
    ref = O.fb_blockwise_f32(G, pos, pairs, T, 1e4, 200, 64)["mean"]
    want = O.log_codes(ref, 8, 1e4)
    codes = ctx.run_fb_codes(pairs, bits=8, core_block_sites=200, flank_sites=64)["codes"]
    d = np.abs(codes.astype(np.int32) - want.astype(np.int32))
    assert d.max() <= 1 and np.count_nonzero(d) <= 1e-3 * d.size
    _case(f"stored states {schedule} budget={budget}: ragged batch")
    jobs = _ragged_jobs(G.shape[0], G.shape[1], np.random.default_rng(29))
    mixed = batch.decode_batch(ctx, jobs, mean_only=True)
    for j, a in zip(jobs, mixed):
        pairs_j = [tuple(map(int, p)) for p in j.pair_array()]
        if j.n_sites == 0 or not pairs_j:
            assert a["mean"].size == 0
            continue
        core_sites = j.core_block_sites if j.core_block_sites > 0 else j.n_sites
        flank = j.flank_sites if j.core_block_sites > 0 else 0
        ref = O.fb_blockwise_f32(np.ascontiguousarray(G[:, j.site_lo:j.site_hi]),
                                 np.ascontiguousarray(pos[j.site_lo:j.site_hi]), pairs_j, T, 1e4, core_sites, flank)
        assert _bits_equal(a["mean"], ref["mean"])


@pytest.mark.parametrize("schedule", ["static", "persistent"])
def test_four_pairs_per_lane_bit_exact(core, cohort_small, flow_field_path, schedule):
    """checkpoint_k 8 with four pairs per lane (256 pairs per wave): means,
    log codes, per-site summaries (which run two pairs per lane) and a
    ragged batch equal the oracle's, also with a partial last tile."""
    from gamma_smc_cu_amd import batch

    G, pos = cohort_small
    pairs = O.all_pairs(G.shape[0]) + [(3, 3)] + O.all_pairs(G.shape[0])[:20]  # 87 pairs: one partial tile
    T = _tables(flow_field_path, 1.25e-8, 1e-8)
    ctx = core.FlowContext(G, pos, 1e4, 1.25e-8, 1e-8, flow_field_path, 0)
    ctx.set_tuning(checkpoint_k=8, pairs_per_lane=4, schedule=schedule, min_waves_per_launch=1)
    _case(f"four pairs per lane {schedule}: blockwise")
    ref = O.fb_blockwise_f32(G, pos, pairs, T, 1e4, 200, 64)["mean"]
    got = ctx.run_fb_blockwise(pairs, core_block_sites=200, flank_sites=64, pair_batch_size=-1, max_streams=2)
    assert _bits_equal(got["mean"], ref)
    _case(f"four pairs per lane {schedule}: u8 codes")
    codes = ctx.run_fb_codes(pairs, bits=8, core_block_sites=200, flank_sites=64)["codes"]
    d = np.abs(codes.astype(np.int32) - O.log_codes(ref, 8, 1e4).astype(np.int32))
    assert d.max() <= 1 and np.count_nonzero(d) <= 1e-3 * d.size
    _case(f"four pairs per lane {schedule}: summary")
    s = ctx.run_fb_summary(pairs, core_block_sites=200, flank_sites=64)
    want = O.site_summary(ref)
    assert _bits_equal(s["site_mean"], want["site_mean"])
    _case(f"four pairs per lane {schedule}: ragged batch")
    jobs = _ragged_jobs(G.shape[0], G.shape[1], np.random.default_rng(41))
    for j, a in zip(jobs, batch.decode_batch(ctx, jobs, mean_only=True)):
        pj = [tuple(map(int, p)) for p in j.pair_array()]
        if j.n_sites == 0 or not pj:
            continue
        core_sites = j.core_block_sites if j.core_block_sites > 0 else j.n_sites
        flank = j.flank_sites if j.core_block_sites > 0 else 0
        r = O.fb_blockwise_f32(np.ascontiguousarray(G[:, j.site_lo:j.site_hi]),
                               np.ascontiguousarray(pos[j.site_lo:j.site_hi]), pj, T, 1e4, core_sites, flank)
        assert _bits_equal(a["mean"], r["mean"])


@pytest.mark.parametrize("schedule", ["static", "persistent"])
def test_checkpoint_phases_bit_exact(core, cohort_small, flow_field_path, schedule):
    """A state budget of about one work item's checkpoints (700 sites / 8 x
    64 pairs x 8 B = 44,800 B) splits checkpoint launches into consecutive
    phases of items that reuse the buffer; every output stays the oracle's,
    bit for bit, with CI and posterior too."""
    G, pos = cohort_small
    pairs = O.all_pairs(G.shape[0]) + [(3, 3)]
    T = _tables(flow_field_path, 1.25e-8, 1e-8)
    ctx = core.FlowContext(G, pos, 1e4, 1.25e-8, 1e-8, flow_field_path, 0)
    ctx.set_tuning(checkpoint_k=8, pairs_per_lane=1, schedule=schedule)
    ctx.set_state_budget(50_000)
    _case(f"checkpoint phases {schedule}: full sequence")
    got = ctx.run_fb(pairs, mean_only=False, return_posterior=True)
    ref = O.fb_window_f32(G, pos, pairs, T, 1e4, ci=True, posterior=True)
    for k, v in ref.items():
        assert _bits_equal(got[k], v), k
    _case(f"checkpoint phases {schedule}: blockwise")
    got = ctx.run_fb_blockwise(pairs, core_block_sites=200, flank_sites=64, pair_batch_size=-1, max_streams=1)
    assert _bits_equal(got["mean"], O.fb_blockwise_f32(G, pos, pairs, T, 1e4, 200, 64)["mean"])


def test_state_budget_below_one_item_raises(core, cohort_small, flow_field_path):
    G, pos = cohort_small
    ctx = core.FlowContext(G, pos, 1e4, 1.25e-8, 1e-8, flow_field_path, 0)
    ctx.set_tuning(checkpoint_k=1)
    ctx.set_state_budget(1000)
    with pytest.raises(MemoryError, match="work item"):
        ctx.run_fb([(1, 0)], mean_only=True)
    ctx.set_state_budget(0)
    assert ctx.run_fb([(1, 0)], mean_only=True)["mean"].shape == (G.shape[1], 1)


def test_device_path_matches_host_path(core, cohort_small, flow_field_path):
    from gamma_smc_cu_amd.device import DeviceArray, Stream, run_fb_device

    G, pos = cohort_small
    pairs = np.array(O.all_pairs(G.shape[0]), np.int32)
    ctx = core.FlowContext(G, pos, 1e4, 1.25e-8, 1e-8, flow_field_path, 0)
    host = ctx.run_fb_blockwise([tuple(p) for p in pairs], core_block_sites=256, flank_sites=64)
    d_pi = DeviceArray.from_host(pairs[:, 0].copy())
    d_pj = DeviceArray.from_host(pairs[:, 1].copy())
    out = DeviceArray((G.shape[1], len(pairs)), np.float32)
    st = Stream()
    _case("device path x3 same stream")
    for _ in range(3):
        run_fb_device(ctx, d_pi, d_pj, len(pairs), 256, 64, 0, out, stream=st)
    st.synchronize()
    assert _bits_equal(out.to_host(), host["mean"])


def test_site_summary_bit_exact_and_shape_invariant(core, cohort_small, flow_field_path):
    """run_fb_summary (per-site mean / min / max over pairs, no per-pair
    output) equals the oracle's fixed-order reduction of the oracle's
    per-pair means, bit for bit, under every launch shape, full-sequence
    and blockwise."""
    G, pos = cohort_small
    mu, rho = O.estimate_scaled_params(G, pos, 1.25e-8, 1e-8, 1e4)
    T = _tables(flow_field_path, mu, rho)
    pairs = O.all_pairs(G.shape[0]) + [(5, 5), (1, 0)]
    ctx = core.FlowContext(G, pos, 1e4, mu, rho, flow_field_path, 0)
    full = O.site_summary(O.fb_window_f32(G, pos, pairs, T, 1e4)["mean"])
    blk = O.site_summary(O.fb_blockwise_f32(G, pos, pairs, T, 1e4, 200, 64)["mean"])
    for tu in SCHEDULES:
        _case(f"summary tuning {tu}")
        ctx.set_tuning(**tu)
        got = ctx.run_fb_summary(pairs)
        assert got["n_pairs"] == len(pairs)
        for k in ("site_mean", "site_min", "site_max"):
            assert _bits_equal(got[k], full[k]), (tu, k)
        got = ctx.run_fb_summary(pairs, core_block_sites=200, flank_sites=64)
        for k in ("site_mean", "site_min", "site_max"):
            assert _bits_equal(got[k], blk[k]), (tu, "blockwise", k)
    _case("summary of no pairs")
    r = ctx.run_fb_summary([])
    assert r["site_mean"].shape == (0,) and r["n_pairs"] == 0 and "site_min" not in r
    with pytest.raises(ValueError, match="out of range"):
        ctx.run_fb_summary([(0, G.shape[0])])


def test_batch_summary_jobs_match_oracle(core, cohort_small, flow_field_path):
    """GSMC_SUMMARY batch jobs: each job's site reductions equal the
    oracle's over that job's pairs and site range."""
    import ctypes

    from gamma_smc_cu_amd.batch import Workspace
    from gamma_smc_cu_amd.device import DeviceArray

    G, pos = cohort_small
    n, S = G.shape
    mu, rho = O.estimate_scaled_params(G, pos, 1.25e-8, 1e-8, 1e4)
    T = _tables(flow_field_path, mu, rho)
    ctx = core.FlowContext(G, pos, 1e4, mu, rho, flow_field_path, 0)
    rng = np.random.default_rng(8)
    specs = []
    for lo, hi, npairs, csz, fl in ((0, S, 130, 0, 0), (50, 420, 7, 128, 30), (300, 301, 64, 0, 0)):
        p = rng.integers(0, n, size=(npairs, 2)).astype(np.int32)
        specs.append((lo, hi, p, csz, fl))
    keep = []
    jobs = (core.Job * len(specs))()
    for k, (lo, hi, p, csz, fl) in enumerate(specs):
        d_pi, d_pj = DeviceArray.from_host(p[:, 0].copy()), DeviceArray.from_host(p[:, 1].copy())
        outs = [DeviceArray((hi - lo,), np.float32) for _ in range(3)]
        keep.append((d_pi, d_pj, outs))
        j = jobs[k]
        j.site_lo, j.site_hi, j.core_block_sites, j.flank_sites = lo, hi, csz, fl
        j.pair_i, j.pair_j, j.n_pairs = d_pi.ptr.value, d_pj.ptr.value, len(p)
        j.site_mean, j.site_min, j.site_max = (o.ptr.value for o in outs)
    _case("batch summary jobs")
    ws = Workspace(ctx)
    ws.launch(jobs, len(specs), core.GSMC_SUMMARY)
    ws.wait()
    for (lo, hi, p, csz, fl), (_, _, outs) in zip(specs, keep):
        sub_G, sub_pos = np.ascontiguousarray(G[:, lo:hi]), pos[lo:hi]
        pl = [tuple(map(int, x)) for x in p]
        if csz:
            ref = O.fb_blockwise_f32(sub_G, sub_pos, pl, T, 1e4, csz, fl)["mean"]
        else:
            ref = O.fb_window_f32(sub_G, sub_pos, pl, T, 1e4)["mean"]
        want = O.site_summary(ref)
        for o, k in zip(outs, ("site_mean", "site_min", "site_max")):
            assert _bits_equal(o.to_host(), want[k]), (lo, hi, k)
    with pytest.raises(ValueError, match="GSMC_SUMMARY"):
        ws.launch(jobs, len(specs), core.GSMC_SUMMARY | core.GSMC_CI)


def _fwd_cohort():
    """Small cohort with the forward-only path's edge cases: duplicate
    positions (gaps that round to 0), gaps of exactly N, and gaps beyond N."""
    from gamma_smc_cu_amd import synthetic

    G, pos = synthetic.cohort(10, 500, 120_000, 19)
    pos = pos.copy()
    pos[100:] += 2500.0          # gap > N = 1000
    pos[200:] += 1000.0 - (pos[200] - pos[199])  # gap exactly N
    pos[300] = pos[299] + 0.25   # rounds to 0: no advance
    pos[301:] += 0.25
    return G, pos


def test_run_fwd_matches_oracle_within_tolerance(core, flow_field_path):
    """run_fwd against the FP32 oracle of the same recursion: the two use
    different (correctly rounded vs 1-ulp) 10^x / log10 / 2^x, so the
    tolerance is stated: |log mean_gpu - log mean_oracle| <= 1e-5; CI bounds
    within 1e-4 relative (the cube in the Wilson-Hilferty bound amplifies
    the last-ulp differences of a near-cancelling base); error against the
    FP64 oracle on FP64 tables (no FP16 rounding) <= 3e-2 max, <= 1e-3
    median: the reference stores this path's tables in FP16 (2^-11 relative
    per node), which dominates."""
    G, pos = _fwd_cohort()
    mu, rho = O.estimate_scaled_params(G, pos, 1.25e-8, 1e-8, 1e4)
    T = _tables(flow_field_path, mu, rho)
    pairs = O.all_pairs(G.shape[0]) + [(4, 4)]
    ctx = core.FlowContext(G, pos, 1e4, mu, rho, flow_field_path, 0)
    _case("run_fwd with CI")
    got = ctx.run_fwd(pairs, mean_only=False)
    ref = O.fwd_only_f32(G, pos, pairs, O.h2_tables(T), 1e4, ci=True)
    err = np.abs(np.log(got["mean"]) - np.log(ref["mean"]))
    print(f"[fwd] max |dlog mean| vs fp32 oracle {err.max():.3e}", flush=True)
    assert err.max() <= 1e-5
    for k in ("lower", "upper"):
        np.testing.assert_allclose(got[k], ref[k], rtol=1e-4, atol=1e-6 * float(ref["mean"].max()))
    assert np.all(got["lower"] <= got["mean"]) and np.all(got["mean"] <= got["upper"])
    r64 = O.fwd_only_f64(G, pos, pairs, O.tables_f64(flow_field_path, 1000, 1e4, mu, rho), 1e4)
    e64 = np.abs(np.log(got["mean"]) - np.log(r64["mean"]))
    print(f"[fwd] vs fp64 oracle: max {e64.max():.3e} median {np.median(e64):.3e}", flush=True)
    assert e64.max() <= 3e-2 and np.median(e64) <= 1e-3
    _case("run_fwd mean only, chunked pair order")
    m_only = ctx.run_fwd(pairs[::-1], mean_only=True)
    assert set(m_only) == {"mean"}
    assert _bits_equal(m_only["mean"][:, ::-1], got["mean"])
    assert ctx.run_fwd([], mean_only=True)["mean"].shape == (G.shape[1], 0)
    with pytest.raises(ValueError, match="out of range"):
        ctx.run_fwd([(0, 99)])


def _nibbles(q):
    return np.stack([q & 0x0F, q >> 4], axis=1).reshape(-1, q.shape[1])


@pytest.mark.parametrize("bits", [8, 4])
def test_log_codes_match_oracle(core, cohort_small, flow_field_path, bits):
    """Log-quantized mean codes equal the oracle's codes of the oracle's
    FP32 means. The GPU's log and the oracle's libm log may differ in the
    last ulp, which can move a value sitting on a code boundary by one
    code: at most one code, in at most 1e-3 of the entries."""
    from gamma_smc_cu_amd._core import dequantize

    G, pos = cohort_small
    mu, rho = O.estimate_scaled_params(G, pos, 1.25e-8, 1e-8, 1e4)
    T = _tables(flow_field_path, mu, rho)
    pairs = O.all_pairs(G.shape[0]) + [(5, 5)]
    for S_used, csz, fl in ((G.shape[1], 200, 64), (G.shape[1] - 1, 0, 0)):
        _case(f"codes bits={bits} S={S_used} core={csz}")
        Gs, ps_ = np.ascontiguousarray(G[:, :S_used]), pos[:S_used]
        ctx = core.FlowContext(Gs, ps_, 1e4, mu, rho, flow_field_path, 0)
        if csz:
            ref = O.fb_blockwise_f32(Gs, ps_, pairs, T, 1e4, csz, fl)["mean"]
        else:
            ref = O.fb_window_f32(Gs, ps_, pairs, T, 1e4)["mean"]
        want = O.log_codes(ref, bits, 1e4)
        got = ctx.run_fb_codes(pairs, bits=bits, core_block_sites=csz, flank_sites=fl)
        assert got["codes"].shape == want.shape and got["bits"] == bits
        a, b = (got["codes"], want) if bits == 8 else (_nibbles(got["codes"]), _nibbles(want))
        d = np.abs(a.astype(np.int32) - b.astype(np.int32))
        assert d.max() <= 1 and np.count_nonzero(d) <= 1e-3 * d.size, (d.max(), np.count_nonzero(d))
        # decoded values sit within half a code of the mean (for means in range)
        deq = dequantize(got["codes"], bits, 1e4)[:S_used]
        half_bin = got["log_max"] / (255.0 if bits == 8 else 15.0)
        inr = (ref >= 1.0) & (ref <= 2e5)
        assert np.all(np.abs(np.log(deq[inr]) - np.log(ref[inr])) <= half_bin * 0.5 + half_bin + 1e-6)
        # 136 pairs (the list repeated): each pair's column is the same in
        # whichever lane and 64-pair group decodes it
        idx = np.arange(136) % len(pairs)
        got_w = ctx.run_fb_codes([pairs[i] for i in idx], bits=bits, core_block_sites=csz, flank_sites=fl)["codes"]
        assert np.array_equal(got_w, got["codes"][:, idx])
    if bits == 4:
        with pytest.raises(ValueError, match="even core_block_sites"):
            ctx.run_fb_codes(pairs, bits=4, core_block_sites=201, flank_sites=10)


@pytest.mark.parametrize("schedule,k", [("static", 16), ("persistent", 16), ("static", 1)])
def test_graph_replays_match_eager(core, cohort_small, flow_field_path, schedule, k):
    """A captured batched launch, replayed several times with other eager
    launches and synchronisations in between, reproduces the eager launch
    bit for bit on every replay (also with stored forward states)."""
    from gamma_smc_cu_amd.batch import Graph, Staged, Workspace
    from gamma_smc_cu_amd.device import Stream

    G, pos = cohort_small
    ctx = core.FlowContext(G, pos, 1e4, 1.25e-8, 1e-8, flow_field_path, 0)
    ctx.set_tuning(checkpoint_k=k, pairs_per_lane=1, schedule=schedule)
    jobs = [j for j in _ragged_jobs(G.shape[0], G.shape[1], np.random.default_rng(5)) if j.n_sites > 0]
    st = Staged(jobs)
    arr, n = st.structs()
    ws = Workspace(ctx)
    _case(f"graph eager reference ({schedule})")
    ws.launch(arr, n, st.flags)
    ws.wait()
    eager = st.d_out.to_host().copy()
    g = Graph(ctx, arr, n, st.flags)
    other = Staged(jobs[:3])
    oarr, on = other.structs()
    s = Stream()
    for r in range(4):
        _case(f"graph replay {r} ({schedule})")
        st.d_out.fill_zero()
        g.launch(s)
        s.synchronize()
        assert _bits_equal(st.d_out.to_host(), eager), r
        ws.launch(oarr, on, other.flags)
        ws.wait()
    g.close()
