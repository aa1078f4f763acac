"""Per-gene statistics: the oracle's restatement of the reference's per-gene
accumulators (``analysis/genome_wide/infer_chromosome.py:157-227``) pinned
by closed forms and an independent float64 loop, and the package's host
logic (gene site ranges, accumulators, summaries) checked against the
oracle. No GPU needed."""

import math

import numpy as np
import pytest

from gamma_smc_cu_amd import genes as GS
from oracle import oracle as O


def _brute(mean, positions, genes, edges):
    """Straight loops in float64: for each gene the sites with start <= pos
    <= end, each pair's mean of max(x, 1) and of log(max(x, 1)) over them,
    then count / sums / minima over pairs and the bin of each pair's mean
    log (edges <= value, less one, clamped)."""
    S, P = mean.shape
    nb = len(edges) - 1
    out = {"count": [], "lin_sum": [], "log_sum": [], "log_sq_sum": [], "min_lin": [], "min_log": [],
           "histogram": []}
    for start, end in genes:
        sites = [s for s in range(S) if start <= positions[s] <= end]
        hist = [0] * nb
        if len(sites) < 2:
            for k, v in (("count", 0), ("lin_sum", 0.0), ("log_sum", 0.0), ("log_sq_sum", 0.0),
                         ("min_lin", math.inf), ("min_log", math.inf)):
                out[k].append(v)
            out["histogram"].append(hist)
            continue
        lins, lgs = [], []
        for p in range(P):
            xs = [max(float(mean[s, p]), 1.0) for s in sites]
            lins.append(sum(xs) / len(xs))
            lgs.append(sum(math.log(x) for x in xs) / len(xs))
        for v in lgs:
            k = sum(1 for e in edges if e <= v)
            hist[min(max(k - 1, 0), nb - 1)] += 1
        out["count"].append(P)
        out["lin_sum"].append(sum(lins))
        out["log_sum"].append(sum(lgs))
        out["log_sq_sum"].append(sum(v * v for v in lgs))
        out["min_lin"].append(min(lins))
        out["min_log"].append(min(lgs))
        out["histogram"].append(hist)
    return {k: np.array(v) for k, v in out.items()}


def _close(a, b, rtol):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    fin = np.isfinite(b)
    assert np.array_equal(np.isfinite(a), fin)
    assert np.allclose(a[fin], b[fin], rtol=rtol, atol=rtol), np.max(np.abs(a[fin] - b[fin]))


def test_gene_accumulators_closed_form():
    """Constant means: every pair's gene values are c and ln c; below 1 the
    floor makes them 1 and 0 (the first bin)."""
    positions = np.arange(10, dtype=np.float64) * 100.0 + 50.0   # 50, 150, ..., 950
    for c in (250.0, 0.3):
        mean = np.full((10, 7), c, np.float32)
        genes = [(150, 450), (50, 50), (2000, 3000), (0, 10_000), (450, 451)]
        acc = O.gene_accumulators(mean, positions, genes)
        x = max(c, 1.0)
        lg = math.log(x)
        width = (math.log(1e6) - math.log(10.0)) / 50
        b = min(max(int(math.floor((lg - math.log(10.0)) / width)) + 1, 1), 50) - 1
        assert acc["count"].tolist() == [7, 0, 0, 7, 0]
        for gi in (0, 3):
            assert acc["lin_sum"][gi] == pytest.approx(7 * x, rel=1e-6)
            assert acc["log_sum"][gi] == pytest.approx(7 * lg, rel=1e-6, abs=1e-7)
            assert acc["log_sq_sum"][gi] == pytest.approx(7 * lg * lg, rel=1e-5, abs=1e-7)
            assert acc["min_lin"][gi] == pytest.approx(x, rel=1e-7)
            assert acc["histogram"][gi].tolist() == [7 if k == b else 0 for k in range(50)]
        for gi in (1, 2, 4):   # 1, 0 and 1 sites: untouched
            assert acc["min_lin"][gi] == math.inf and acc["histogram"][gi].sum() == 0


def test_gene_accumulators_against_float64_loops():
    """Random means, overlapping / nested / edge-touching genes, two pair
    chunks accumulated: the oracle (float32 as the reference computes) and
    plain float64 loops agree to float32 rounding, and the histograms
    exactly. Gene bounds sit exactly on site positions, so an exclusive end
    or start would drop sites and fail."""
    rng = np.random.default_rng(11)
    S, P = 60, 9
    positions = np.cumsum(rng.integers(20, 400, S)).astype(np.float64)
    mean = np.exp(rng.normal(8.0, 2.0, (S, P))).astype(np.float32)
    mean[rng.random((S, P)) < 0.05] = 0.5      # below the floor
    p = positions
    genes = [(p[0], p[-1]), (p[3], p[9]), (p[5], p[7]), (p[9], p[20]), (p[30], p[30]), (p[40] + 1, p[41] - 1),
             (p[44], p[45]), (p[10] - 5, p[12] + 5), (p[-1] + 10, p[-1] + 20)]
    edges = O.GENE_BIN_EDGES
    acc = O.gene_accumulators(mean[:, :4], positions, genes, edges)
    acc = O.gene_accumulators(mean[:, 4:], positions, genes, edges, acc)
    ref = _brute(mean, positions, genes, edges)
    assert acc["count"].tolist() == ref["count"].tolist()
    for k in ("lin_sum", "log_sum", "log_sq_sum", "min_lin", "min_log"):
        _close(acc[k], ref[k], 2e-6)
    assert np.array_equal(acc["histogram"], ref["histogram"])
    assert ref["count"][4] == 0 and ref["count"][5] == 0 and ref["count"][6] == P


def test_gene_site_ranges_match_oracle_indices():
    """The package's [lo, hi) ranges are the oracle's index sets, also with
    repeated positions, reversed and out-of-range genes."""
    rng = np.random.default_rng(3)
    positions = np.sort(rng.integers(0, 5000, 400)).astype(np.float64)   # repeats included
    genes = [(int(a), int(a + w)) for a, w in zip(rng.integers(-100, 5100, 60), rng.integers(-50, 800, 60))]
    genes += [(positions[7], positions[7]), (positions[0], positions[-1]), (6000, 7000), (-10, -1)]
    r = GS.gene_site_ranges(positions, genes)
    for (lo, hi), idx in zip(r, O.gene_site_indices(positions, genes)):
        assert list(range(lo, hi)) == idx.tolist()
    rows = [("ENSG1", "A", 10, 20), ("ENSG2", "B", 15, 30)]
    assert np.array_equal(GS.gene_site_ranges(positions, rows), GS.gene_site_ranges(positions, [(10, 20), (15, 30)]))
    with pytest.raises(ValueError, match="sorted"):
        GS.gene_site_ranges(positions[::-1], genes)
    with pytest.raises(ValueError, match="start, end"):
        GS.gene_site_ranges(positions, [(1, 2, 3)])


def test_accumulators_and_summary():
    acc = GS.new_accumulators(3)
    assert acc["histogram"].shape == (3, 50) and np.array_equal(acc["bin_edges"], O.GENE_BIN_EDGES)
    assert np.all(np.isinf(acc["min_lin"])) and acc["count"].dtype == np.int64
    acc["count"][:] = [0, 4, 2]
    acc["log_sum"][:] = [0.0, 4 * math.log(100.0), 2 * math.log(7.0)]
    acc["lin_sum"][:] = [0.0, 400.0, 14.0]
    acc["min_lin"][1:] = [90.0, 7.0]
    s = GS.summarize(acc)
    assert np.isnan(s["geom_mean_tmrca"][0]) and np.isnan(s["arith_mean_tmrca"][0]) and np.isnan(s["min_tmrca"][0])
    assert s["geom_mean_tmrca"][1:] == pytest.approx([100.0, 7.0])
    assert s["arith_mean_tmrca"][1:] == pytest.approx([100.0, 7.0])
    custom = GS.new_accumulators(2, np.linspace(0.0, 1.0, 5))
    assert custom["histogram"].shape == (2, 4)


def test_run_fb_gene_stats_checks_accumulators_before_the_library():
    """The library adds into the accumulator arrays in place, so arrays of
    another dtype, shape or layout are refused before any native call."""
    from gamma_smc_cu_amd import _core

    ctx = _core.FlowContext.__new__(_core.FlowContext)   # no device needed: the check comes first
    ranges = [(0, 5), (1, 3)]
    for key, bad in (("lin_sum", lambda a: a.astype(np.float32)), ("count", lambda a: a[:1]),
                     ("histogram", lambda a: np.asfortranarray(np.zeros((2, 50), np.int64))),
                     ("min_log", lambda a: np.broadcast_to(a, a.shape))):
        acc = GS.new_accumulators(2)
        acc[key] = bad(acc[key])
        with pytest.raises(ValueError, match=key):
            ctx.run_fb_gene_stats([(0, 1)], ranges, accumulators=acc)
    with pytest.raises(ValueError, match="histogram"):
        ctx.run_fb_gene_stats([(0, 1)], ranges, accumulators=GS.new_accumulators(2, np.linspace(0, 1, 4)))


def test_write_results_matches_the_pipeline_layout(tmp_path):
    """CSV columns, two-decimal means, blanks for genes without pairs, and
    the NPZ fields of infer_chromosome.py:239-281 (read back without
    pickle)."""
    import csv

    genes = [("ENSG1", "AAA", 100, 900), ("ENSG2", "BBB", 1000, 1001)]
    acc = GS.new_accumulators(2)
    acc["count"][0] = 4
    acc["lin_sum"][0] = 4 * 1234.5
    acc["log_sum"][0] = 4 * math.log(1000.0)
    acc["min_lin"][0] = 800.25
    acc["n_sites_per_gene"] = np.array([12, 1], np.int32)
    GS.write_results(acc, genes, tmp_path / "pop.csv", tmp_path / "pop.npz", n_haplotypes=8, n_pairs_total=4)
    rows = list(csv.reader(open(tmp_path / "pop.csv")))
    assert rows[0] == ["gene_id", "gene_name", "start", "end", "geom_mean_tmrca", "arith_mean_tmrca", "min_tmrca",
                       "n_pairs", "n_sites"]
    assert rows[1] == ["ENSG1", "AAA", "100", "900", "1000.00", "1234.50", "800.25", "4", "12"]
    assert rows[2] == ["ENSG2", "BBB", "1000", "1001", "", "", "", "0", "1"]
    z = np.load(tmp_path / "pop.npz")
    assert z["gene_id"].tolist() == ["ENSG1", "ENSG2"] and z["start"].tolist() == [100, 1000]
    assert np.array_equal(z["histogram"], acc["histogram"]) and np.array_equal(z["bin_edges"], O.GENE_BIN_EDGES)
    assert int(z["n_haplotypes"]) == 8 and int(z["n_pairs_total"]) == 4 and z["count"].tolist() == [4, 0]
    with pytest.raises(ValueError, match="gene_id"):
        GS.write_results(acc, [(100, 900)], tmp_path / "x.csv")
