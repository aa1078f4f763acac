"""Per-gene statistics of pairwise TMRCA posterior means.

The reference's production pipeline (``analysis/genome_wide/
infer_chromosome.py:140-227``) decodes a population's pairs in chunks with
``infer_blockwise``, brings each chunk's ``(n_sites, n_pairs)`` means to the
host and reduces them there, per gene, into running accumulators (written
out as ``:259-276``). :func:`gene_stats` returns the same accumulators while
the means are never stored (``FlowContext.run_fb_gene_stats``: the decode's
backward kernel sums each pair's values per gene piece in place of writing
rows, and one more kernel reduces the pieces per gene).

Per gene ``g`` with ``n_g >= 2`` sites ``(positions >= start) & (positions
<= end)`` and each pair: ``lin`` = mean over the gene's sites of
``max(mean, 1)`` and ``lg`` = mean of its natural log; over the pairs,
``count`` (pairs), ``lin_sum``, ``log_sum``, ``log_sq_sum`` (sum of
``lg**2``), ``min_lin``, ``min_log`` and ``histogram`` (``lg`` binned by
``numpy.digitize(lg, bin_edges) - 1``, clipped to the bins). Genes of
fewer than two sites keep their initial values, as in the reference.

The reference sums in float32 (numpy ``mean`` / ``sum`` of float32 arrays
and a float32 ``log``); the device sums in float64 over the float32 means
and the hardware log2 scaled by ln 2, so the two agree to float32 rounding,
and a per-pair ``lg`` within that distance of a bin edge may land in the
neighbouring bin.
"""

from __future__ import annotations

import numpy as np

# infer_chromosome.py:52-58: 50 natural-log bins from ln(10) to ln(1e6).
DEFAULT_BIN_EDGES = np.linspace(np.log(10.0), np.log(1_000_000.0), 51)


def new_accumulators(n_genes, bin_edges=None):
    """Empty running accumulators for ``n_genes`` genes
    (``infer_chromosome.py:157-165``)."""
    edges = np.asarray(DEFAULT_BIN_EDGES if bin_edges is None else bin_edges, dtype=np.float64)
    return {
        "count": np.zeros(n_genes, np.int64),
        "lin_sum": np.zeros(n_genes, np.float64),
        "log_sum": np.zeros(n_genes, np.float64),
        "log_sq_sum": np.zeros(n_genes, np.float64),
        "min_lin": np.full(n_genes, np.inf, np.float64),
        "min_log": np.full(n_genes, np.inf, np.float64),
        "histogram": np.zeros((n_genes, edges.size - 1), np.int64),
        "bin_edges": edges.copy(),
    }


def _gene_bounds(genes):
    """(start, end) bp of each gene: 2-tuples, or the reference's
    ``(gene_id, gene_name, start, end)`` rows (``load_genes``, ``:84-92``)."""
    out = []
    for g in genes:
        g = tuple(g)
        if len(g) == 4:
            out.append((int(g[2]), int(g[3])))
        elif len(g) == 2:
            out.append((int(g[0]), int(g[1])))
        else:
            raise ValueError("genes must be (start, end) or (gene_id, gene_name, start, end)")
    return out


def gene_site_ranges(positions, genes):
    """Site ranges ``[lo, hi)`` of each gene over sorted ``positions``: the
    sites with ``start <= position <= end`` (``compute_gene_site_indices``,
    ``infer_chromosome.py:101-112``, which returns them as index arrays)."""
    pos = np.asarray(positions, dtype=np.float64)
    if pos.size > 1 and np.any(np.diff(pos) < 0):
        raise ValueError("positions must be sorted in increasing order")
    b = _gene_bounds(genes)
    if not b:
        return np.zeros((0, 2), np.int64)
    st = np.array([x[0] for x in b], dtype=np.float64)
    en = np.array([x[1] for x in b], dtype=np.float64)
    lo = np.searchsorted(pos, st, side="left")
    hi = np.maximum(np.searchsorted(pos, en, side="right"), lo)
    return np.stack([lo, hi], axis=1).astype(np.int64)


def gene_stats(G_or_ts, positions=None, genes=(), mu=1.25e-8, rho=1e-8, Ne=10000, pairs=None,
               flow_field_path=None, core_block_sites="auto", flank_sites=2048, auto_estimate_theta=True,
               bin_edges=None, accumulators=None, pair_batch_size=-1):
    """Per-gene accumulators of the posterior means over ``pairs`` (default:
    all pairs ``i < j`` in the reference pipeline's order, ``make_pairs``,
    ``infer_chromosome.py:97-98``). Inputs are prepared as
    ``infer_blockwise`` prepares them (segregating sites only, auto theta,
    ``core_block_sites="auto"`` sized from free device memory), and genes
    are matched against the kept sites' positions, as the reference matches
    them against ``result["positions"]``. Pass an earlier result as
    ``accumulators`` to add more pairs to it; ``pair_batch_size`` caps the
    pairs per launch (-1: as many as device memory holds).

    Returns the accumulators and ``n_sites_per_gene`` (int32), ``n_pairs``
    (this call) and ``positions`` (the kept sites)."""
    from . import _core
    from .infer import (_coerce_inputs, _estimate_scaled_params, _filter_segregating, _normalize_pairs,
                        _query_free_gpu_bytes, _recommend_core_block_size, _resolve_flow_field_path,
                        _DEFAULT_BLOCK_SITES)

    if flank_sites < 0:
        raise ValueError("flank_sites must be non-negative.")
    G, positions = _coerce_inputs(G_or_ts, positions)
    G, positions, _ = _filter_segregating(G, positions)
    n_haps, n_sites = G.shape
    if pairs is None:
        pairs = [(i, j) for i in range(n_haps) for j in range(i + 1, n_haps)]
    pairs = _normalize_pairs(pairs)
    ranges = gene_site_ranges(positions, genes)
    acc = new_accumulators(ranges.shape[0], bin_edges) if accumulators is None else accumulators
    if core_block_sites in ("auto", None):
        free_bytes = _query_free_gpu_bytes()
        if free_bytes is None:
            core_block_sites = _DEFAULT_BLOCK_SITES
        else:
            core_block_sites, _ = _recommend_core_block_size(n_sites, max(len(pairs), 1), True, flank_sites,
                                                             free_bytes)
    if flank_sites == 0 and core_block_sites < n_sites:
        raise ValueError("gene_stats: flank_sites=0 with core_block_sites < n_sites restarts every block "
                         "from the prior; pass flank_sites >= 256 or core_block_sites >= n_sites.")
    if pairs and n_sites and ranges.shape[0]:
        flow_field_path = _resolve_flow_field_path(flow_field_path)
        if auto_estimate_theta:
            kmu, krho = _estimate_scaled_params(G, positions, mu, rho, Ne)
        else:
            kmu, krho = float(mu), float(rho)
        ctx = _core.FlowContext(G, positions, float(Ne), kmu, krho, flow_field_path, 0)
        try:
            ctx.run_fb_gene_stats(pairs, ranges, core_block_sites=int(core_block_sites), flank_sites=flank_sites,
                                  bin_edges=acc["bin_edges"], accumulators=acc, pair_batch_size=pair_batch_size)
        finally:
            ctx.close()
    acc["n_sites_per_gene"] = (ranges[:, 1] - ranges[:, 0]).astype(np.int32)
    acc["n_pairs"] = len(pairs)
    acc["positions"] = positions
    return acc


def summarize(acc):
    """Per-gene geometric and arithmetic means of the pairs' values, NaN for
    genes without pairs (``infer_chromosome.py:234-237``)."""
    count = acc["count"]
    with np.errstate(divide="ignore", invalid="ignore"):
        geom = np.where(count > 0, np.exp(acc["log_sum"] / count), np.nan)
        arith = np.where(count > 0, acc["lin_sum"] / count, np.nan)
    return {"geom_mean_tmrca": geom, "arith_mean_tmrca": arith,
            "min_tmrca": np.where(np.isfinite(acc["min_lin"]), acc["min_lin"], np.nan)}


def write_results(acc, genes, csv_path=None, npz_path=None, n_haplotypes=None, n_pairs_total=None):
    """Write the reference pipeline's per-gene outputs
    (``infer_chromosome.py:239-281``): a CSV with ``gene_id, gene_name,
    start, end, geom_mean_tmrca, arith_mean_tmrca, min_tmrca, n_pairs,
    n_sites`` (means to two decimals, blank where a gene has no pairs) and a
    compressed NPZ of the raw accumulators for offline re-aggregation.
    ``genes`` are ``(gene_id, gene_name, start, end)`` rows; ``n_pairs_total``
    defaults to the largest per-gene pair count."""
    import csv

    rows = [tuple(g) for g in genes]
    if any(len(g) != 4 for g in rows):
        raise ValueError("genes must be (gene_id, gene_name, start, end) rows")
    summ = summarize(acc)
    n_sites = acc.get("n_sites_per_gene", np.zeros(len(rows), np.int32))
    fmt = lambda v: f"{v:.2f}" if np.isfinite(v) else ""
    if csv_path is not None:
        with open(csv_path, "w", newline="") as fh:
            w = csv.writer(fh)
            w.writerow(["gene_id", "gene_name", "start", "end", "geom_mean_tmrca", "arith_mean_tmrca", "min_tmrca",
                        "n_pairs", "n_sites"])
            for gi, (gid, name, start, end) in enumerate(rows):
                w.writerow([gid, name, start, end, fmt(summ["geom_mean_tmrca"][gi]),
                            fmt(summ["arith_mean_tmrca"][gi]), fmt(acc["min_lin"][gi]), int(acc["count"][gi]),
                            int(n_sites[gi])])
    if npz_path is not None:
        np.savez_compressed(
            npz_path,
            gene_id=np.array([g[0] for g in rows]), gene_name=np.array([g[1] for g in rows]),
            start=np.array([g[2] for g in rows], dtype=np.int64), end=np.array([g[3] for g in rows], dtype=np.int64),
            count=acc["count"], lin_sum=acc["lin_sum"], log_sum=acc["log_sum"], log_sq_sum=acc["log_sq_sum"],
            min_lin=acc["min_lin"], min_log=acc["min_log"], histogram=acc["histogram"], bin_edges=acc["bin_edges"],
            n_sites_per_gene=np.asarray(n_sites, np.int32),
            n_haplotypes=np.int64(-1 if n_haplotypes is None else n_haplotypes),
            n_pairs_total=np.int64(n_pairs_total if n_pairs_total is not None
                                   else (int(acc["count"].max()) if len(rows) else 0)))


__all__ = ["DEFAULT_BIN_EDGES", "new_accumulators", "gene_site_ranges", "gene_stats", "summarize", "write_results"]
