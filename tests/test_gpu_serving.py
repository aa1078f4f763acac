"""GPU tests of the serving loop: continuous batching over a resident
bitpacked cohort. Every query's result must be bit-identical whether it ran
alone (one query per launch) or batched with others. Both runs must match
the FP32 oracle. Admission must respect the arena budget."""

import importlib

import numpy as np
import pytest

import gamma_smc_cu_amd
from gamma_smc_cu_amd import serving, synthetic
from oracle import oracle as O

pytestmark = pytest.mark.gpu
infer_mod = importlib.import_module("gamma_smc_cu_amd.infer")

N_HAPS, N_SITES = 48, 4000


def _case(msg):
    print(f"[case] {msg}", flush=True)


@pytest.fixture(scope="module")
def resident(flow_field_path):
    packed, pos = synthetic.resident_cohort(N_HAPS, N_SITES, 600_000, 31, chunk=256)
    mu, rho = infer_mod._estimate_scaled_params_packed(packed, N_HAPS, pos, 1.25e-8, 1e-8, 1e4)
    ctx = gamma_smc_cu_amd.FlowContext(packed, pos, 1e4, mu, rho, flow_field_path, 0, packed=True,
                                       n_sites=N_SITES)
    return ctx, packed, pos, mu, rho


def _queries(seed, n):
    rng = np.random.default_rng(seed)
    qs = []
    for _ in range(n):
        m = int(rng.integers(2, 12))
        h = np.sort(rng.choice(N_HAPS, size=m, replace=False))
        S = int(rng.integers(40, 1500))
        lo = int(rng.integers(0, N_SITES - S + 1))
        qs.append(serving.Query(h, lo, lo + S, core_block_sites=400, flank_sites=60))
    return qs


def _copy(q):
    return serving.Query(q.haplotypes, q.site_lo, q.site_hi, q.core_block_sites, q.flank_sites)


def _alone(ctx, qs):
    srv = serving.Server(ctx, arena_bytes=64 << 20, n_slots=1, ck_bytes_per_launch=16 << 20,
                         max_items_per_launch=4096, max_queries_per_launch=1)
    out = []
    for q in qs:
        c = _copy(q)
        srv.submit(c)
        srv.drain()
        out.append(srv.result(c))
        srv.release(c)
    assert srv.launches == len(qs)
    return out


def _bits_equal(a, b):
    return a.shape == b.shape and np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_burst_batches_and_overlaps(resident):
    ctx = resident[0]
    qs = _queries(1, 40)
    _case("serving burst: 40 queries, 2 slots, 24 items per launch")
    srv = serving.Server(ctx, arena_bytes=256 << 20, n_slots=2, ck_bytes_per_launch=16 << 20,
                         max_items_per_launch=24, max_queries_per_launch=4096)
    got = {}

    def keep(q):
        got[q.qid] = srv.result(q)
        srv.release(q)

    serving.run_open_loop(srv, qs, np.zeros(len(qs)), keep)
    assert len(got) == len(qs)
    assert srv.launches < len(qs), "queries were not batched"
    assert srv.max_in_flight == 2, "launches on the two slots did not overlap"
    assert srv.arena.used == 0 and srv.arena.peak <= srv.arena.size
    assert all(q.finished >= q.arrival for q in qs)
    _case("serving burst: alone reruns")
    alone = _alone(ctx, qs)
    for q, a in zip(qs, alone):
        assert _bits_equal(got[q.qid], a), q


def test_staggered_tight_memory_matches_alone_and_oracle(resident, flow_field_path):
    ctx, packed, pos, mu, rho = resident
    qs = _queries(2, 30)
    biggest = max(serving._align(q.n_sites * len(serving.subset_pairs(q.haplotypes)) * 4) for q in qs)
    arena = 3 * (biggest + (64 << 10))
    _case(f"serving staggered: 30 queries, arena {arena} B")
    srv = serving.Server(ctx, arena_bytes=arena, n_slots=2, ck_bytes_per_launch=16 << 20, max_items_per_launch=4096)
    got = {}

    def keep(q):
        got[q.qid] = srv.result(q)
        srv.release(q)

    serving.run_open_loop(srv, qs, np.linspace(0.0, 0.03, len(qs)), keep)
    assert len(got) == len(qs) and srv.arena.peak <= srv.arena.size and srv.arena.used == 0
    alone = _alone(ctx, qs)
    for q, a in zip(qs, alone):
        assert _bits_equal(got[q.qid], a), q
    _case("serving staggered: oracle")
    G = O.unpack(packed, N_HAPS, N_SITES)
    T = O.tables_f32(flow_field_path, 1000, 1e4, mu, rho)
    for q in qs[:4]:
        pairs = [tuple(map(int, p)) for p in q.pairs]
        ref = O.fb_blockwise_f32(np.ascontiguousarray(G[:, q.site_lo:q.site_hi]), pos[q.site_lo:q.site_hi], pairs,
                                 T, 1e4, q.core_block_sites, q.flank_sites)
        assert _bits_equal(got[q.qid], ref["mean"])


def test_rejects_queries_that_cannot_fit(resident):
    ctx = resident[0]
    srv = serving.Server(ctx, arena_bytes=1 << 16, n_slots=1, ck_bytes_per_launch=1 << 20, max_items_per_launch=64)
    with pytest.raises(ValueError, match="HBM budget"):
        srv.submit(serving.Query(np.arange(40), 0, N_SITES))
    with pytest.raises(ValueError, match="outside"):
        srv.submit(serving.Query([0, 1], 10, N_SITES + 1))
    with pytest.raises(ValueError, match="distinct"):
        srv.submit(serving.Query([3, 3], 0, 10))
    assert srv.idle()


def test_code_queries_match_float_queries(resident):
    """Queries asking for uint8 / uint4 codes batch only with queries of
    the same encoding, and their codes are those of the float queries'
    means (to one code at a log ulp boundary)."""
    ctx = resident[0]
    base = _queries(3, 12)
    mixed = []
    for k, q in enumerate(base):
        mixed.append(serving.Query(q.haplotypes, q.site_lo, q.site_hi, 400, 60, bits=(0, 8, 4)[k % 3]))
    _case("serving mixed encodings")
    srv = serving.Server(ctx, arena_bytes=64 << 20, n_slots=2, ck_bytes_per_launch=16 << 20, max_items_per_launch=4096)
    got = {}

    def keep(q):
        got[q.qid] = srv.result(q)
        srv.release(q)

    serving.run_open_loop(srv, mixed, np.zeros(len(mixed)), keep)
    floats = _alone(ctx, [q for q in base])
    for q, f in zip(mixed, floats):
        r = got[q.qid]
        if q.bits == 0:
            assert _bits_equal(r, f)
            continue
        want = O.log_codes(f, q.bits, 1e4)
        assert r.shape == want.shape and r.dtype == np.uint8
        a, b = (r, want) if q.bits == 8 else (_nib(r), _nib(want))
        d = np.abs(a.astype(np.int32) - b.astype(np.int32))
        assert d.max() <= 1 and np.count_nonzero(d) <= max(1, 1e-3 * d.size)


def _nib(q):
    return np.stack([q & 0x0F, q >> 4], axis=1).reshape(-1, q.shape[1])
