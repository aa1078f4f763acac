"""Host logic of the serving loop and the synthetic workload generators
(no GPU): the arena allocator, subset pair order, query streams, the ragged
job generator, the resident cohort generator and the packed auto-theta."""

import importlib

import numpy as np
import pytest

from gamma_smc_cu_amd import serving, synthetic
from oracle import oracle as O

infer_mod = importlib.import_module("gamma_smc_cu_amd.infer")


def test_arena_matches_bitmap_model():
    rng = np.random.default_rng(0)
    a = serving.Arena(64 * 256)
    owner = np.zeros(64, bool)  # granule occupancy
    live = []
    for _ in range(2000):
        if live and rng.random() < 0.45:
            off, n = live.pop(int(rng.integers(len(live))))
            a.release(off, n)
            owner[off // 256:(off + serving._align(n)) // 256] = False
        else:
            n = int(rng.integers(1, 9 * 256))
            off = a.alloc(n)
            g = serving._align(n) // 256
            free_runs = [s for s in range(64 - g + 1) if not owner[s:s + g].any()]
            if off is None:
                assert not free_runs, "first fit missed a free run"
            else:
                assert off % 256 == 0 and off // 256 == free_runs[0]
                owner[off // 256:off // 256 + g] = True
                live.append((off, n))
        assert a.used == owner.sum() * 256
        # free list sorted, disjoint, merged, and exactly the unowned granules
        runs = a.free
        assert all(runs[i][0] + runs[i][1] < runs[i + 1][0] for i in range(len(runs) - 1))
        free = np.zeros(64, bool)
        for off, n in runs:
            free[off // 256:(off + n) // 256] = True
        assert np.array_equal(free, ~owner)
    assert a.peak <= a.size


def test_subset_pairs_follow_all_pairs_order():
    h = [17, 3, 40, 8]
    p = serving.subset_pairs(h)
    hs = sorted(h)
    want = [(hs[i], hs[j]) for i, j in O.all_pairs(len(hs))]
    assert [tuple(map(int, x)) for x in p] == want
    assert p.dtype == np.int32


def test_poisson_queries_deterministic_and_bounded():
    q1, t1 = serving.poisson_queries(200, 50.0, 6404, 1_000_000, 13)
    q2, t2 = serving.poisson_queries(200, 50.0, 6404, 1_000_000, 13)
    assert np.array_equal(t1, t2) and np.all(np.diff(t1) > 0)
    for a, b in zip(q1, q2):
        assert np.array_equal(a.haplotypes, b.haplotypes) and (a.site_lo, a.site_hi) == (b.site_lo, b.site_hi)
        m = len(a.haplotypes)
        assert 4 <= m <= 200 and len(np.unique(a.haplotypes)) == m
        S = a.site_hi - a.site_lo
        assert 0 <= a.site_lo and a.site_hi <= 1_000_000
        assert round(50_000 / 150) <= S <= round(5_000_000 / 150)
    # mean inter-arrival close to 1/rate
    assert abs(np.diff(np.concatenate([[0.0], t1])).mean() - 1 / 50.0) < 0.3 / 50.0


def test_ragged_jobs_shapes():
    jobs = synthetic.ragged_jobs(2000, 262_144)
    assert len(jobs) == 64
    for lo, hi, p in jobs:
        assert 0 <= lo < hi <= 262_144 and 2_000 <= hi - lo <= 200_000
        assert 100 <= len(p) <= 20_000
        assert (p[:, 0] > p[:, 1]).all() and (p[:, 1] >= 0).all() and (p[:, 0] < 2000).all()
        k = p[:, 0].astype(np.int64) * (p[:, 0] - 1) // 2 + p[:, 1]
        assert np.unique(k).size == k.size
        assert all(O.ij_to_pair(int(a), int(b)) == int(x) for (a, b), x in zip(p[:20], k[:20]))


def test_resident_cohort_exact_counts_and_thread_invariance():
    p1, pos1 = synthetic.resident_cohort(30, 4000, 800_000, 9, chunk=256, threads=1)
    p3, pos3 = synthetic.resident_cohort(30, 4000, 800_000, 9, chunk=256, threads=3)
    assert np.array_equal(p1, p3) and np.array_equal(pos1, pos3)
    G = O.unpack(p1, 30, 4000)
    rng = np.random.default_rng(9)
    want_pos = synthetic._positions(rng, 4000, 800_000)
    want_k = synthetic._counts(rng, 30, 4000)
    assert np.array_equal(pos1, want_pos)
    assert np.array_equal(G.sum(axis=0), want_k)
    assert np.array_equal(O.bitpack(G), p1)  # tail bits zero, layout as bitpack


@pytest.mark.parametrize("n", [40, 41, 2])
def test_packed_theta_matches_unpacked(n):
    G, pos = synthetic.cohort(n, 900, 300_000, 5)
    a = infer_mod._estimate_scaled_params(G, pos, 1.25e-8, 1e-8, 1e4)
    b = infer_mod._estimate_scaled_params_packed(O.bitpack(G), n, pos, 1.25e-8, 1e-8, 1e4)
    assert a == b
