"""GPU: per-gene statistics (``gsmc_run_fb_gene_stats``) against the oracle's
restatement of the reference's per-gene accumulators
(``analysis/genome_wide/infer_chromosome.py:157-227``) applied to the
oracle's FP32 posterior means (which the device decode matches bit for bit).

Tolerance: the reference, and so the oracle, reduces in float32 (numpy
``mean`` / ``sum`` and ``log`` of float32 arrays); the device sums the same
float32 means in float64 and takes the log as the hardware log2 times
ln 2. Counts agree exactly; sums and minima to a relative 1e-5 (absolute
1e-6 near zero); a pair's value moves to the neighbouring histogram bin only
where the oracle's value lies within 1e-5 of a bin edge.
"""

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _case(msg):
    print(f"[case] {msg}", flush=True)


@pytest.fixture(scope="module")
def core():
    from gamma_smc_cu_amd import _core

    assert _core.device_count() >= 1
    return _core


def _assert_matches(got, mean, positions, genes, edges=O.GENE_BIN_EDGES):
    want = O.gene_accumulators(mean, positions, genes, edges)
    assert got["count"].tolist() == want["count"].tolist()
    for k in ("lin_sum", "log_sum", "log_sq_sum", "min_lin", "min_log"):
        a, b = got[k], want[k]
        assert np.array_equal(np.isinf(a), np.isinf(b)), k
        fin = np.isfinite(b)
        assert np.allclose(a[fin], b[fin], rtol=1e-5, atol=1e-6), (k, np.max(np.abs(a[fin] - b[fin])))
    vals = O.gene_pair_values(mean, positions, genes)
    e = np.asarray(edges)
    for g, v in enumerate(vals):
        h, w = got["histogram"][g], want["histogram"][g]
        assert h.sum() == w.sum(), g
        if v is None:
            assert h.sum() == 0
            continue
        lg = v[1].astype(np.float64)
        near = int(np.count_nonzero(np.min(np.abs(lg[:, None] - e[None, :]), axis=1) <= 1e-5))
        assert np.abs(h - w).sum() <= 2 * near, (g, np.abs(h - w).sum(), near)


def _mixed_genes(pos):
    """Whole sequence, block-crossing, nested and duplicate genes, a
    one-site gene, a gene between two sites, genes past the end."""
    p = pos
    return [(p[0], p[-1]), (p[150], p[260]), (p[190], p[210]), (p[195], p[205]), (p[150], p[260]),
            (p[300], p[300]), (p[400] + 0.25, p[400] + 0.5), (p[-1] + 10, p[-1] + 100), (p[650], p[-1] + 1000),
            (p[0] - 500, p[3]), (p[201], p[202])]


@pytest.mark.parametrize("csz,fl", [(200, 64), (0, 0)])
def test_gene_stats_match_oracle(core, cohort_small, flow_field_path, csz, fl):
    from gamma_smc_cu_amd.genes import gene_site_ranges

    G, pos = cohort_small
    mu, rho = O.estimate_scaled_params(G, pos, 1.25e-8, 1e-8, 1e4)
    T = O.tables_f32(flow_field_path, 1000, 1e4, mu, rho)
    pairs = O.all_pairs(G.shape[0]) + [(5, 5)]
    genes = _mixed_genes(pos)
    _case(f"gene stats core={csz} flank={fl} P={len(pairs)}")
    ctx = core.FlowContext(G, pos, 1e4, mu, rho, flow_field_path, 0)
    got = ctx.run_fb_gene_stats(pairs, gene_site_ranges(pos, genes), core_block_sites=csz, flank_sites=fl)
    if csz:
        ref = O.fb_blockwise_f32(G, pos, pairs, T, 1e4, csz, fl)["mean"]
    else:
        ref = O.fb_window_f32(G, pos, pairs, T, 1e4)["mean"]
    _assert_matches(got, ref, pos, genes)
    assert got["count"][5] == 0 and got["count"][6] == 0 and got["count"][7] == 0 and got["count"][10] == len(pairs)


def test_gene_stats_accumulate_over_pair_chunks(core, cohort_small, flow_field_path):
    """Two calls over halves of the pairs, the second adding to the first's
    accumulators, and one call in launches of 20 pairs give the one-launch
    result: per-pair values do not depend on the chunk, so counts, minima
    and histograms are equal and the sums differ only by the order of the
    float64 additions."""
    from gamma_smc_cu_amd.genes import gene_site_ranges

    G, pos = cohort_small
    pairs = O.all_pairs(G.shape[0])
    r = gene_site_ranges(pos, _mixed_genes(pos))
    ctx = core.FlowContext(G, pos, 1e4, 1.25e-8, 1e-8, flow_field_path, 0)
    _case("gene stats, one call")
    one = ctx.run_fb_gene_stats(pairs, r, core_block_sites=200, flank_sites=64)
    _case("gene stats, two chunks")
    two = ctx.run_fb_gene_stats(pairs[:30], r, core_block_sites=200, flank_sites=64)
    two = ctx.run_fb_gene_stats(pairs[30:], r, core_block_sites=200, flank_sites=64, accumulators=two)
    _case("gene stats, launches of 20 pairs")
    four = ctx.run_fb_gene_stats(pairs, r, core_block_sites=200, flank_sites=64, pair_batch_size=20)
    for other in (two, four):
        for k in ("count", "min_lin", "min_log", "histogram"):
            assert np.array_equal(one[k], other[k]), k
        for k in ("lin_sum", "log_sum", "log_sq_sum"):
            assert np.allclose(one[k], other[k], rtol=1e-12, atol=0), k
    with pytest.raises(ValueError, match="pair_batch_size"):
        ctx.run_fb_gene_stats(pairs, r, pair_batch_size=0)


def test_gene_stats_same_under_every_tuning(core, cohort_small, flow_field_path):
    """The gene epilogue sums each piece in decreasing site order whatever
    the checkpoint interval, and gene launches run one pair per lane on the
    static schedule whatever the tuning: every field is bit-identical."""
    from gamma_smc_cu_amd.genes import gene_site_ranges

    G, pos = cohort_small
    pairs = O.all_pairs(G.shape[0])
    r = gene_site_ranges(pos, _mixed_genes(pos))
    ctx = core.FlowContext(G, pos, 1e4, 1.25e-8, 1e-8, flow_field_path, 0)
    base = None
    for k, np_, sch in ((8, 1, "static"), (16, 2, "persistent"), (1, 1, "persistent"), (8, 4, "dynamic")):
        _case(f"gene stats tuning K={k} np={np_} {sch}")
        ctx.set_tuning(checkpoint_k=k, pairs_per_lane=np_, schedule=sch)
        got = ctx.run_fb_gene_stats(pairs, r, core_block_sites=200, flank_sites=64)
        if base is None:
            base = got
            continue
        for f in ("count", "lin_sum", "log_sum", "log_sq_sum", "min_lin", "min_log", "histogram"):
            assert np.array_equal(got[f], base[f]), (k, np_, sch, f)


def test_gene_stats_long_genes(core, flow_field_path):
    """9,000 sites in three block cores: genes spanning one, two and three
    cores (their pieces split at the cores' boundaries), genes ending on a
    core boundary, short genes at both ends."""
    from gamma_smc_cu_amd import synthetic
    from gamma_smc_cu_amd.genes import gene_site_ranges

    G, pos = synthetic.cohort(8, 9000, 9000 * 150, 21)
    mu, rho = O.estimate_scaled_params(G, pos, 1.25e-8, 1e-8, 1e4)
    T = O.tables_f32(flow_field_path, 1000, 1e4, mu, rho)
    pairs = O.all_pairs(8)
    p = pos
    genes = [(p[0], p[-1]), (p[100], p[8500]), (p[2999], p[3000]), (p[4000], p[4100]), (p[4095], p[5999]),
             (p[10], p[12]), (p[8990], p[8999])]
    _case("gene stats, 8 x 9000, core 3000")
    ctx = core.FlowContext(G, pos, 1e4, mu, rho, flow_field_path, 0)
    got = ctx.run_fb_gene_stats(pairs, gene_site_ranges(pos, genes), core_block_sites=3000, flank_sites=300)
    ref = O.fb_blockwise_f32(G, pos, pairs, T, 1e4, 3000, 300)["mean"]
    _assert_matches(got, ref, pos, genes)


def test_gene_stats_api_matches_infer_blockwise(cohort_small):
    """gene_stats() prepares inputs as infer_blockwise does: monomorphic
    sites dropped, auto theta, genes matched to the kept positions. Its
    result equals the oracle's accumulators over infer_blockwise's means."""
    import gamma_smc_cu_amd as g

    G, pos = cohort_small
    rng = np.random.default_rng(2)
    extra = np.sort(rng.choice(np.arange(1, G.shape[1]), 25, replace=False))
    G2 = np.insert(G, extra, 0, axis=1)
    pos2 = np.insert(pos, extra, (pos[extra - 1] + pos[extra]) / 2.0)
    pairs = [(i, j) for i in range(G.shape[0]) for j in range(i + 1, G.shape[0])]
    genes = [("g1", "A", pos[10], pos[400]), ("g2", "B", pos2[extra[3]], pos[600]), ("g3", "C", pos[5], pos[6])]
    _case("gene_stats API")
    got = g.gene_stats(G2, pos2, genes, pairs=pairs, core_block_sites=200, flank_sites=64)
    res = g.infer_blockwise(G2, pos2, pairs=pairs, core_block_sites=200, flank_sites=64)
    kept = np.asarray(res["positions"])
    assert np.array_equal(np.asarray(got["positions"]), kept)
    bounds = [(x[2], x[3]) for x in genes]
    _assert_matches(got, np.asarray(res["mean"]), kept, bounds)
    assert got["n_sites_per_gene"].tolist() == [len(i) for i in O.gene_site_indices(kept, bounds)]
    assert got["n_pairs"] == len(pairs)


def test_gene_stats_rejects_bad_arguments(core, cohort_small, flow_field_path):
    G, pos = cohort_small
    ctx = core.FlowContext(G, pos, 1e4, 1.25e-8, 1e-8, flow_field_path, 0)
    pairs = [(0, 1)]
    S = G.shape[1]
    with pytest.raises(ValueError, match="site range"):
        ctx.run_fb_gene_stats(pairs, [(5, S + 1)])
    with pytest.raises(ValueError, match="site range"):
        ctx.run_fb_gene_stats(pairs, [(7, 6)])
    with pytest.raises(ValueError, match="increase"):
        ctx.run_fb_gene_stats(pairs, [(0, 5)], bin_edges=[0.0, 2.0, 1.0])
    with pytest.raises(ValueError, match="n_bins"):
        ctx.run_fb_gene_stats(pairs, [(0, 5)], bin_edges=np.linspace(0, 1, 2000))
    with pytest.raises(ValueError, match="out of range"):
        ctx.run_fb_gene_stats([(0, G.shape[0])], [(0, 5)])
    acc = ctx.run_fb_gene_stats([], [(0, 5)])
    assert acc["count"].tolist() == [0]
