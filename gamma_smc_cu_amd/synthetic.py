"""Synthetic cohorts for the decoder's workloads.

Every cohort is a set of biallelic segregating sites on ``n`` haplotypes:

* positions: ``S`` distinct integer coordinates drawn uniformly on
  ``[0, span)`` and sorted;
* derived-allele count per site: ``k`` drawn with probability proportional
  to ``1 / k`` on ``1 .. n - 1`` (the neutral site-frequency spectrum);
* carriers: ``k`` haplotypes chosen uniformly without replacement.

Draw order for a given ``seed`` (``numpy.random.default_rng(seed)``):
positions, then all ``k``, then carriers in site chunks of ``chunk`` sites
(one ``random((chunk, n))`` matrix per chunk, carriers = the ``k`` smallest
keys). The generator can emit the uint8 matrix ``G[n, S]`` or the bitpacked
rows directly (bit ``s & 63`` of word ``s >> 6``), the latter without ever
holding ``G`` for large cohorts.
"""

from __future__ import annotations

import numpy as np


def _positions(rng, S, span):
    if S > span:
        raise ValueError(f"cannot place {S} distinct sites on a span of {span}")
    pos = rng.choice(int(span), size=int(S), replace=False)
    pos.sort()
    return pos.astype(np.float64)


def _counts(rng, n, S):
    k = np.arange(1, n, dtype=np.int64)
    p = 1.0 / k
    p /= p.sum()
    return rng.choice(k, size=int(S), p=p)


def _carrier_chunks(rng, n, counts, chunk):
    S = counts.shape[0]
    for s0 in range(0, S, chunk):
        s1 = min(S, s0 + chunk)
        keys = rng.random((s1 - s0, n))
        rank = keys.argsort(axis=1).argsort(axis=1)
        yield s0, s1, (rank < counts[s0:s1, None])  # (chunk, n) bool


def cohort(n, S, span, seed, chunk=8192):
    """``(G uint8 (n, S), positions float64 (S,))`` for a 1/k-spectrum cohort."""
    rng = np.random.default_rng(seed)
    pos = _positions(rng, S, span)
    counts = _counts(rng, n, S)
    G = np.zeros((n, S), np.uint8)
    for s0, s1, carr in _carrier_chunks(rng, n, counts, chunk):
        G[:, s0:s1] = carr.T
    return G, pos


def cohort_packed(n, S, span, seed, chunk=8192):
    """Same cohort as :func:`cohort`, returned as ``(packed uint64 (n, W), positions)``."""
    if chunk % 64:
        raise ValueError("chunk must be a multiple of 64")
    rng = np.random.default_rng(seed)
    pos = _positions(rng, S, span)
    counts = _counts(rng, n, S)
    W = (S + 63) // 64
    packed = np.zeros((n, W), np.uint64)
    weights = np.uint64(1) << np.arange(64, dtype=np.uint64)
    for s0, s1, carr in _carrier_chunks(rng, n, counts, chunk):
        m = s1 - s0
        mw = (m + 63) // 64
        bits = np.zeros((n, mw * 64), np.uint64)
        bits[:, :m] = carr.T
        packed[:, s0 // 64:s0 // 64 + mw] = (bits.reshape(n, mw, 64) * weights).sum(axis=2, dtype=np.uint64)
    return packed, pos


# Workload shapes -------------------------------------------------------

CHROMOSOME = dict(n=356, S=402_853, span=51_000_000, seed=7)      # 63,190 pairs, blockwise mean_only
BIOBANK = dict(n=1000, S=50_000, span=7_500_000, seed=17)         # 499,500 pairs
SERVING = dict(n=6404, S=1_000_000, span=150_000_000, seed=13)    # resident cohort for the request loop
RAGGED_COHORT = dict(n=2000, seed=11, bp_per_site=150)             # batched ragged jobs


def ragged_jobs(n_haps, n_sites, n_jobs=64, seed=11, sites=(2_000, 200_000), pairs=(100, 20_000)):
    """Ragged decode jobs over a cohort of ``n_haps`` x ``n_sites``.

    Job k draws, in order, from ``numpy.random.default_rng([seed, 1])``: its
    site count ``exp(U(log lo, log hi))`` (rounded, at most ``n_sites``), its
    pair count likewise, the first site uniformly among the starts that fit,
    and its pairs as distinct indices into the all-pairs order (pair
    ``p = i(i-1)/2 + j``, ``i > j``). Returns a list of
    ``(site_lo, site_hi, pairs int32 (P, 2))``."""
    rng = np.random.default_rng([seed, 1])
    total = n_haps * (n_haps - 1) // 2
    out = []
    for _ in range(n_jobs):
        S = int(min(n_sites, round(np.exp(rng.uniform(np.log(sites[0]), np.log(sites[1]))))))
        P = int(min(total, round(np.exp(rng.uniform(np.log(pairs[0]), np.log(pairs[1]))))))
        lo = int(rng.integers(0, n_sites - S + 1))
        idx = np.sort(rng.choice(total, size=P, replace=False))
        i = np.floor((1.0 + np.sqrt(1.0 + 8.0 * idx.astype(np.float64))) / 2.0).astype(np.int64)
        i -= (i * (i - 1) // 2 > idx)
        i += ((i + 1) * i // 2 <= idx)
        j = idx - i * (i - 1) // 2
        out.append((lo, lo + S, np.stack([i, j], 1).astype(np.int32)))
    return out


def genes(span, n_genes=800, seed=29, length=(1_000, 200_000)):
    """Synthetic gene intervals on ``[0, span)`` bp for per-gene statistics:
    gene k draws, from ``numpy.random.default_rng([seed, 2])``, a length
    ``exp(U(log lo, log hi))`` bp and a start uniform among those that fit.
    Genes may overlap, as annotated genes do. Returns ``(start, end)``
    integer pairs (``end`` inclusive), sorted by start."""
    rng = np.random.default_rng([seed, 2])
    ln = np.exp(rng.uniform(np.log(length[0]), np.log(length[1]), n_genes)).astype(np.int64)
    st = (rng.random(n_genes) * np.maximum(span - ln, 1)).astype(np.int64)
    order = np.argsort(st, kind="stable")
    return [(int(st[k]), int(st[k] + ln[k] - 1)) for k in order]


def resident_cohort(n, S, span, seed, chunk=2048, threads=None):
    """Bitpacked 1/k-spectrum cohort for large ``n`` (the serving cohort):
    ``(packed uint64 (n, W), positions float64 (S,))``.

    Positions and counts are drawn as in :func:`cohort` from
    ``default_rng(seed)``. Carriers of the sites in chunk ``c`` (``chunk``
    sites each) come from their own stream ``default_rng([seed, 2, c])``:
    one ``random((chunk, n), float32)`` key matrix, and a site's carriers
    are its ``k`` smallest keys. The chunks are independent, so they are
    drawn on a thread pool and the result does not depend on ``threads``.
    """
    from concurrent.futures import ThreadPoolExecutor

    if chunk % 64:
        raise ValueError("chunk must be a multiple of 64")
    rng = np.random.default_rng(seed)
    pos = _positions(rng, S, span)
    counts = _counts(rng, n, S)
    W = (S + 63) // 64
    packed = np.zeros((n, W), np.uint64)
    cols = np.arange(n)[None, :]

    def one(c):
        s0, s1 = c * chunk, min(S, (c + 1) * chunk)
        keys = np.random.default_rng([seed, 2, c]).random((s1 - s0, n), dtype=np.float32)
        order = keys.argsort(axis=1)
        carr = np.zeros(keys.shape, bool)
        np.put_along_axis(carr, order, cols < counts[s0:s1, None], axis=1)
        m = s1 - s0
        mw = (m + 63) // 64
        bits = np.zeros((n, mw * 64), bool)
        bits[:, :m] = carr.T
        packed[:, s0 // 64:s0 // 64 + mw] = np.packbits(bits, axis=1, bitorder="little").view("<u8")

    n_chunks = (S + chunk - 1) // chunk
    with ThreadPoolExecutor(max_workers=threads or min(16, len(__import__("os").sched_getaffinity(0)))) as ex:
        list(ex.map(one, range(n_chunks)))
    return packed, pos
