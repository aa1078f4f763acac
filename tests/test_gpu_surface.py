"""GPU tests of the rest of the reference's ``_core`` surface on this path
(``bindings.cpp:3618-3645,3704``): ``gamma_smc_flow_cached_fb`` /
``gamma_smc_flow_cached_fwd``, ``get_device_count`` / ``get_device`` /
``set_device`` and ``FlowContext.run_fwd_states``, each against the FP32
oracle; ``max_streams`` output slots; and the fixed checkpoint buffer of a
reserved workspace."""

import numpy as np
import pytest

import gamma_smc_cu_amd
from gamma_smc_cu_amd import _core, synthetic
from gamma_smc_cu_amd.batch import Job, Staged, Workspace
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _case(msg):
    print(f"[case] {msg}", flush=True)


def _eq_bits(a, b):
    return a.shape == b.shape and np.array_equal(np.ascontiguousarray(a).view(np.uint32),
                                                 np.ascontiguousarray(b).view(np.uint32))


@pytest.fixture(scope="module")
def shifted_cohort():
    # positions start past 3 Mb, so the first site's gap spans thousands of
    # full N-bp stretches (the prior-table start)
    G, pos = synthetic.cohort(14, 900, 500_000, 23)
    return G, pos + 3_000_000.0


def test_device_functions():
    _case("device functions")
    assert _core.get_device_count() >= 1
    _core.set_device(0)
    assert _core.get_device() == 0


def test_run_fwd_states_bit_exact(shifted_cohort, flow_field_path):
    G, pos = shifted_cohort
    mu, rho = O.estimate_scaled_params(G, pos, 1.25e-8, 1e-8, 1e4)
    pairs = O.all_pairs(G.shape[0])
    _case(f"run_fwd_states {len(pairs)} pairs x {G.shape[1]} sites")
    ctx = gamma_smc_cu_amd.FlowContext(G, pos, 1e4, mu, rho, flow_field_path, 0)
    r = ctx.run_fwd_states(pairs)
    T = O.tables_f32(flow_field_path, 1000, 1e4, mu, rho)
    for p, pair in enumerate(pairs):
        m, c = O.forward_states_f32(G, pos, pair, T)
        assert _eq_bits(r["mean_log10"][:, p], m), pair
        assert _eq_bits(r["cv_log10"][:, p], c), pair
    assert ctx.run_fwd_states([])["mean_log10"].shape == (G.shape[1], 0)


def test_gamma_smc_flow_cached_fb_matches_oracle(shifted_cohort, flow_field_path):
    G, pos = shifted_cohort
    pairs = O.all_pairs(G.shape[0])[:50]
    mu, rho = 1.25e-8, 1e-8  # the given rates: no heterozygosity estimate on this entry point
    _case("gamma_smc_flow_cached_fb")
    r = gamma_smc_cu_amd.gamma_smc_flow_cached_fb(G, pos, pairs, 1e4, mu, rho, flow_field_path)
    assert set(r) == {"mean", "lower", "upper"}  # mean_only defaults to False, as in the reference
    T = O.tables_f32(flow_field_path, 1000, 1e4, mu, rho)
    ref = O.fb_window_f32(G, pos, pairs, T, 1e4, ci=True)
    for k in ("mean", "lower", "upper"):
        assert _eq_bits(r[k], ref[k]), k
    r1 = _core.gamma_smc_flow_cached_fb(G, pos, pairs, 1e4, mu, rho, flow_field_path, True, 0)
    assert set(r1) == {"mean"} and _eq_bits(r1["mean"], ref["mean"])


def test_gamma_smc_flow_cached_fwd_matches_run_fwd(shifted_cohort, flow_field_path):
    G, pos = shifted_cohort
    pairs = O.all_pairs(G.shape[0])[:20]
    _case("gamma_smc_flow_cached_fwd")
    a = gamma_smc_cu_amd.gamma_smc_flow_cached_fwd(G, pos, pairs, 1e4, 1.25e-8, 1e-8, flow_field_path)
    ctx = gamma_smc_cu_amd.FlowContext(G, pos, 1e4, 1.25e-8, 1e-8, flow_field_path, 0)
    assert _eq_bits(a["mean"], ctx.run_fwd(pairs)["mean"])


@pytest.mark.parametrize("k,schedule", [(8, "static"), (1, "persistent")])
def test_max_streams_slots_give_identical_results(shifted_cohort, flow_field_path, k, schedule):
    """More output slots (streams) than groups, fewer, and one: the same
    rows, bit for bit, and the oracle's."""
    G, pos = shifted_cohort
    mu, rho = O.estimate_scaled_params(G, pos, 1.25e-8, 1e-8, 1e4)
    pairs = O.all_pairs(G.shape[0])
    ctx = gamma_smc_cu_amd.FlowContext(G, pos, 1e4, mu, rho, flow_field_path, 0)
    ctx.set_tuning(checkpoint_k=k, schedule=schedule, min_waves_per_launch=1)  # one block per launch group
    outs = []
    for ms in (1, 2, 3, 5):
        _case(f"max_streams={ms} K={k} {schedule}")
        outs.append(ctx.run_fb_blockwise(pairs, core_block_sites=128, flank_sites=40, pair_batch_size=30,
                                         max_streams=ms)["mean"])
    T = O.tables_f32(flow_field_path, 1000, 1e4, mu, rho)
    ref = O.fb_blockwise_f32(G, pos, pairs, T, 1e4, 128, 40)["mean"]
    for o in outs:
        assert _eq_bits(o, ref)


def test_reserved_workspace_keeps_its_checkpoint_buffer(flow_field_path):
    """Stored forward states, persistent schedule: a launch mixing one long
    job with many short ones plans state slots of the longest item for every
    resident wave. On a reserved workspace that launch runs with fewer
    resident waves instead of growing the buffer, and each job's rows equal
    its rows decoded alone."""
    G, pos = synthetic.cohort(24, 6000, 1_500_000, 5)
    mu, rho = O.estimate_scaled_params(G, pos, 1.25e-8, 1e-8, 1e4)
    ctx = gamma_smc_cu_amd.FlowContext(G, pos, 1e4, mu, rho, flow_field_path, 0)
    ctx.set_tuning(checkpoint_k=1, schedule="persistent")
    allp = O.all_pairs(24)
    jobs = [Job(0, 6000, allp[:64], core_block_sites=4096, flank_sites=256)]
    jobs += [Job(100 * i, 100 * i + 300, allp[i:i + 3], core_block_sites=256, flank_sites=32) for i in range(30)]
    need = sum(ctx.checkpoint_bytes(j.n_sites, j.core_block_sites, j.flank_sites, len(j.pairs)) for j in jobs)
    ws = Workspace(ctx)
    ws.reserve(need, 4096, 64)
    assert ws.checkpoint_buffer_bytes() == need
    _case("reserved mixed launch")
    st = Staged(jobs)
    arr, n = st.structs()
    ws.launch(arr, n, st.flags)
    ws.wait()
    mix = st.results()
    assert ws.checkpoint_buffer_bytes() == need
    alone = Workspace(ctx)
    for k in (0, 1, 29):
        arr, n = st.structs([k])
        alone.launch(arr, n, st.flags)
        alone.wait()
    res = st.results()
    for k in (0, 1, 29):
        assert _eq_bits(res[k]["mean"], mix[k]["mean"]), k


@pytest.mark.parametrize("k,schedule", [(1, "persistent"), (8, "persistent"), (8, "static")])
def test_workspace_reused_across_streams(flow_field_path, k, schedule):
    """One workspace, launches with different metadata on two streams and
    the null stream in turn, queued without host waits between them: each
    launch first waits (on the host) for the workspace's previous launch
    before its metadata upload and queue-counter reset, so no kernel reads
    metadata or counters another launch is rewriting. Every job's rows equal
    the oracle's."""
    from gamma_smc_cu_amd.device import Stream

    G, pos = synthetic.cohort(16, 1500, 450_000, 9)
    mu, rho = O.estimate_scaled_params(G, pos, 1.25e-8, 1e-8, 1e4)
    ctx = gamma_smc_cu_amd.FlowContext(G, pos, 1e4, mu, rho, flow_field_path, 0)
    ctx.set_tuning(checkpoint_k=k, schedule=schedule)
    rng = np.random.default_rng(3)
    allp = O.all_pairs(16)
    jobs = []
    for _ in range(9):
        lo = int(rng.integers(0, 1000))
        hi = int(rng.integers(lo + 50, 1500))
        sel = rng.choice(len(allp), size=int(rng.integers(5, 90)), replace=False)
        jobs.append(Job(lo, hi, [allp[s] for s in sel], core_block_sites=int(rng.choice([0, 200])), flank_sites=40))
    st = Staged(jobs)
    ws = Workspace(ctx)
    s1, s2 = Stream(), Stream()
    streams = [s1, None, s2]
    _case(f"workspace across streams K={k} {schedule}")
    for i in range(len(jobs)):
        arr, n = st.structs([i])
        ws.launch(arr, n, st.flags, streams[i % 3])
    ws.wait()
    s1.synchronize()
    s2.synchronize()
    T = O.tables_f32(flow_field_path, 1000, 1e4, mu, rho)
    for j, r in zip(jobs, st.results()):
        pairs = [tuple(map(int, p)) for p in j.pair_array()]
        core_sites = j.core_block_sites if j.core_block_sites > 0 else j.n_sites
        flank = j.flank_sites if j.core_block_sites > 0 else 0
        ref = O.fb_blockwise_f32(np.ascontiguousarray(G[:, j.site_lo:j.site_hi]), pos[j.site_lo:j.site_hi], pairs, T,
                                 1e4, core_sites, flank)
        assert _eq_bits(r["mean"], ref["mean"])
