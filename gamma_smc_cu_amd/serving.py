"""Serving loop: TMRCA queries against a cohort resident in HBM.

A query asks for the posterior mean of every pair of a haplotype subset over
a region ``[site_lo, site_hi)`` of the resident cohort. The region is decoded
as one sequence, the way ``infer_blockwise`` decodes a cohort restricted to
those sites: cores of ``core_block_sites`` with ``flank_sites`` flanks, the
first site's gap measured from position -1. Pairs are ``(h[a], h[b])``,
``a > b``, over the sorted subset ``h``, in the all-pairs order
``(1,0), (2,0), (2,1), ...``. The result is ``mean[n_sites, n_pairs]``.

Continuous batching.
- The server keeps up to ``n_slots`` launches in flight. Each slot has its own
  stream and launch workspace.
- Whenever a slot is free it admits waiting queries, oldest first, while they
  fit the HBM budget:
  - outputs go into a device arena allocated once, so the request path
    never allocates device memory;
  - checkpoints and the launch's pair ids go into the slot's own buffers,
    reserved up front for the per-launch caps on checkpoint bytes and work
    items (the pair ids of a whole launch cross PCIe in one pinned copy).
- All queries admitted together run as one batched launch. Queries arriving
  while launches run wait for the next free slot and for memory. Launches on
  different slots overlap on the device.
- Lanes never mix pairs, so a query's result is bit-identical alone and in
  any mix.

Times are host ``time.perf_counter()`` seconds. A query's ``finished`` time is
its launch's completion event, mapped to the host clock through a reference
event recorded at start-up.
"""

from __future__ import annotations

import collections
import ctypes
import math
import time
from dataclasses import dataclass, field

import numpy as np

from . import _core
from .batch import Workspace
from .device import DeviceArray, Event, Stream, synchronize

_GRANULE = 256


def _align(n):
    return (int(n) + _GRANULE - 1) // _GRANULE * _GRANULE


class Arena:
    """First-fit allocator of byte ranges of ``[0, size)`` in 256-byte
    granules; host bookkeeping only. Freed neighbours are merged."""

    def __init__(self, size):
        self.size = int(size) // _GRANULE * _GRANULE
        self.free = [(0, self.size)] if self.size else []
        self.used = 0
        self.peak = 0

    def alloc(self, nbytes):
        n = _align(max(int(nbytes), 1))
        for k, (off, length) in enumerate(self.free):
            if length >= n:
                if length == n:
                    del self.free[k]
                else:
                    self.free[k] = (off + n, length - n)
                self.used += n
                self.peak = max(self.peak, self.used)
                return off
        return None

    def release(self, off, nbytes):
        n = _align(max(int(nbytes), 1))
        k = 0
        while k < len(self.free) and self.free[k][0] < off:
            k += 1
        self.free.insert(k, (off, n))
        # merge with the right, then the left neighbour
        if k + 1 < len(self.free) and off + n == self.free[k + 1][0]:
            self.free[k] = (off, n + self.free[k + 1][1])
            del self.free[k + 1]
        if k > 0 and self.free[k - 1][0] + self.free[k - 1][1] == off:
            self.free[k - 1] = (self.free[k - 1][0], self.free[k - 1][1] + self.free[k][1])
            del self.free[k]
        self.used -= n


def subset_pairs(haplotypes):
    """All pairs of a haplotype subset: ``(h[a], h[b])`` for ``a > b`` over
    the sorted subset, in the all-pairs order."""
    h = np.unique(np.asarray(haplotypes, dtype=np.int64))
    a, b = np.tril_indices(h.shape[0], k=-1)
    return np.stack([h[a], h[b]], axis=1).astype(np.int32)


@dataclass
class Query:
    haplotypes: object
    site_lo: int
    site_hi: int
    core_block_sites: int = 32_768
    flank_sites: int = 2048
    bits: int = 0              # output: 0 float32 means; 8 / 4 log-quantized codes
    qid: int = -1
    arrival: float = math.nan
    launched: float = math.nan
    finished: float = math.nan
    pairs: np.ndarray = field(default=None, repr=False)
    _off: int = -1
    _bytes: int = 0
    _ck: int = 0
    _items: int = 0

    @property
    def n_sites(self):
        return int(self.site_hi) - int(self.site_lo)

    @property
    def n_pairs(self):
        return 0 if self.pairs is None else int(self.pairs.shape[0])

    @property
    def pair_sites(self):
        return self.n_sites * self.n_pairs

    @property
    def latency(self):
        return self.finished - self.arrival

    @property
    def out_shape(self):
        rows = self.n_sites if self.bits != 4 else (self.n_sites + 1) // 2
        return rows, self.n_pairs

    @property
    def out_dtype(self):
        return np.float32 if self.bits == 0 else np.uint8



def _n_blocks(n_sites, core, flank):
    if core <= 0 or n_sites <= core:
        return 1
    return -(-n_sites // core)


class _PinnedHost:
    def __init__(self):
        self.ptr = ctypes.c_void_p()
        self.nbytes = 0

    def ensure(self, nbytes):
        if nbytes <= self.nbytes:
            return
        self.close()
        n = max(int(nbytes), 1 << 20)
        _core.check(_core.lib().gsmc_host_alloc(ctypes.byref(self.ptr), n))
        self.nbytes = n

    def view(self, offset, count):
        return np.ctypeslib.as_array((ctypes.c_int32 * count).from_address(self.ptr.value + offset))

    def close(self):
        if self.ptr.value:
            _core.lib().gsmc_host_free(self.ptr)
        self.ptr = ctypes.c_void_p()
        self.nbytes = 0

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class _Slot:
    def __init__(self, ctx, ck_bytes, max_items, max_jobs):
        self.ws = Workspace(ctx)
        _core.check(_core.lib().gsmc_batch_reserve(self.ws.handle, int(ck_bytes), int(max_items), int(max_jobs)))
        self.stream = Stream()
        self.start = Event()
        self.done = Event()
        # pair ids of one launch: [pair_i of every query | pair_j of every query]
        self.max_pairs = 64 * int(max_items)
        self.stage = _PinnedHost()
        self.stage.ensure(8 * self.max_pairs)
        self.d_pairs = DeviceArray((2 * self.max_pairs,), np.int32)
        self.queries = []

    @property
    def busy(self):
        return bool(self.queries)


class Server:
    """Continuous-batching decoder over a resident cohort.

    ``arena_bytes``: device memory for the outputs of queries in flight and
    not yet released. ``ck_bytes_per_launch`` /
    ``max_items_per_launch`` / ``max_queries_per_launch``: caps on one
    launch (checkpoints, work items, queries). The caps bound the memory each
    slot reserves and a launch's duration. ``max_queries_per_launch=1``
    with ``n_slots=1`` is the one-query-per-launch mode.
    """

    def __init__(self, ctx, arena_bytes, n_slots=2, ck_bytes_per_launch=8 << 30, max_items_per_launch=1 << 17,
                 max_queries_per_launch=4096, clock=time.perf_counter):
        if n_slots < 1:
            raise ValueError("n_slots must be >= 1")
        self.ctx = ctx
        self.clock = clock
        self.arena = Arena(arena_bytes)
        self.d_arena = DeviceArray((max(self.arena.size, _GRANULE),), np.uint8)
        self.ck_cap = int(ck_bytes_per_launch)
        self.item_cap = int(max_items_per_launch)
        self.query_cap = int(max_queries_per_launch)
        self.slots = [_Slot(ctx, self.ck_cap, self.item_cap, self.query_cap) for _ in range(n_slots)]
        self.pending = collections.deque()
        self.completed = {}
        self.launches = 0
        self.max_in_flight = 0  # most launches seen running at once
        self.launch_log = []    # per finished launch: (queries, pair-sites, work items, device ms)
        self._next_id = 0
        synchronize()
        self._ev0 = Event()
        t_a = clock()
        self._ev0.record(self.slots[0].stream)
        self.slots[0].stream.synchronize()
        self._t0 = 0.5 * (t_a + clock())

    # -- requests ---------------------------------------------------------
    def submit(self, q, arrival=None):
        """Queue a query; returns its id. Raises ValueError for a query that
        is malformed or can never fit the budget."""
        S, n = self.ctx.n_sites, self.ctx.n_haps
        if not (0 <= int(q.site_lo) < int(q.site_hi) <= S):
            raise ValueError(f"region [{q.site_lo}, {q.site_hi}) outside [0, {S})")
        h = np.asarray(q.haplotypes, dtype=np.int64)
        if h.size < 2 or np.unique(h).size != h.size or h.min() < 0 or h.max() >= n:
            raise ValueError("haplotypes must be at least two distinct ids in [0, n_haps)")
        if q.flank_sites < 0:
            raise ValueError("flank_sites must be non-negative.")
        if q.bits not in (0, 8, 4):
            raise ValueError("bits must be 0 (float means), 8 or 4")
        if q.bits == 4 and q.core_block_sites > 0 and q.n_sites > q.core_block_sites and q.core_block_sites % 2:
            raise ValueError("uint4 output needs an even core_block_sites")
        q.pairs = subset_pairs(h)
        rows, cols = q.out_shape
        q._bytes = rows * cols * np.dtype(q.out_dtype).itemsize
        q._ck = self.ctx.checkpoint_bytes(q.n_sites, q.core_block_sites, q.flank_sites, q.n_pairs)
        q._items = -(-q.n_pairs // 64) * _n_blocks(q.n_sites, q.core_block_sites, q.flank_sites)
        if _align(q._bytes) > self.arena.size or q._ck > self.ck_cap or q._items > self.item_cap:
            raise ValueError(f"query of {q.n_pairs} pairs x {q.n_sites} sites exceeds the HBM budget")
        q.qid = self._next_id
        self._next_id += 1
        q.arrival = self.clock() if arrival is None else float(arrival)
        self.pending.append(q)
        return q.qid

    def poll(self):
        """Retire finished launches and admit waiting queries into free
        slots, without blocking. Returns the queries that finished."""
        finished = []
        flag = ctypes.c_int()
        for slot in self.slots:
            if not slot.busy:
                continue
            _core.check(_core.lib().gsmc_event_query(slot.done.handle, ctypes.byref(flag)))
            if flag.value:
                t = self._t0 + self._ev0.elapsed_ms(slot.done) / 1e3
                self.launch_log.append((len(slot.queries), sum(q.pair_sites for q in slot.queries),
                                        sum(q._items for q in slot.queries), slot.start.elapsed_ms(slot.done)))
                for q in slot.queries:
                    q.finished = t
                    self.completed[q.qid] = q
                    finished.append(q)
                slot.queries = []
        for slot in self.slots:
            if not self.pending:
                break
            if not slot.busy:
                batch = self._admit()
                if not batch:
                    break
                self._launch(slot, batch)
        return finished

    def idle(self):
        return not self.pending and not any(s.busy for s in self.slots)

    def drain(self, sleep=50e-6):
        out = []
        while not self.idle():
            got = self.poll()
            out += got
            if not got:
                time.sleep(sleep)
        return out

    def result(self, q):
        """Copy a finished query's output to the host: float32 means
        ``(n_sites, n_pairs)``, or uint8 codes (``bits`` 8: same shape; 4:
        ``(ceil(n_sites / 2), n_pairs)``, see ``FlowContext.run_fb_codes``)."""
        if q.qid not in self.completed:
            raise ValueError("query has not finished")
        out = np.empty(q.out_shape, q.out_dtype)
        src = self.d_arena.offset(q._off)
        _core.check(_core.lib().gsmc_memcpy_dtoh(ctypes.c_void_p(out.ctypes.data), src, out.nbytes))
        return out

    def release(self, q):
        """Return a finished query's memory to the arena."""
        if self.completed.pop(q.qid, None) is not None:
            self.arena.release(q._off, q._bytes)
            q._off = -1

    # -- internals --------------------------------------------------------
    def _admit(self):
        batch, ck, items = [], 0, 0
        while self.pending and len(batch) < self.query_cap:
            q = self.pending[0]
            if batch and (ck + q._ck > self.ck_cap or items + q._items > self.item_cap or q.bits != batch[0].bits):
                break
            off = self.arena.alloc(q._bytes)
            if off is None:
                break  # oldest first: wait until memory is released
            q._off = off
            batch.append(self.pending.popleft())
            ck += q._ck
            items += q._items
        return batch

    def _launch(self, slot, batch):
        total = sum(q.n_pairs for q in batch)
        pi = slot.stage.view(0, total)
        pj = slot.stage.view(4 * total, total)
        jobs = (_core.Job * len(batch))()
        base = self.d_arena.ptr.value
        dp = slot.d_pairs.ptr.value
        at = 0
        for k, q in enumerate(batch):
            P = q.n_pairs
            pi[at:at + P] = q.pairs[:, 0]
            pj[at:at + P] = q.pairs[:, 1]
            j = jobs[k]
            j.site_lo, j.site_hi = int(q.site_lo), int(q.site_hi)
            j.core_block_sites, j.flank_sites = int(q.core_block_sites), int(q.flank_sites)
            j.pair_i, j.pair_j = dp + 4 * at, dp + 4 * (total + at)
            j.n_pairs = P
            j.mean = base + q._off
            at += P
        _core.check(_core.lib().gsmc_memcpy_htod_async(slot.d_pairs.ptr, slot.stage.ptr, 8 * total,
                                                       slot.stream.handle))
        flags = {0: 0, 8: _core.GSMC_Q8, 4: _core.GSMC_Q4}[batch[0].bits]
        slot.start.record(slot.stream)
        slot.ws.launch(jobs, len(batch), flags, slot.stream)
        slot.done.record(slot.stream)
        now = self.clock()
        for q in batch:
            q.launched = now
        slot.queries = batch
        self.launches += 1
        self.max_in_flight = max(self.max_in_flight, sum(s.busy for s in self.slots))


def poisson_queries(n_queries, rate, n_haps, n_sites, seed, subset=(4, 200), region_bp=(50_000, 5_000_000),
                    bp_per_site=150.0, core_block_sites=32_768, flank_sites=2048):
    """Open-loop query stream: Poisson arrivals at ``rate`` per second.

    Query k draws, from ``numpy.random.default_rng([seed, 3])`` in order:
    - its inter-arrival gap ``Exp(1/rate)``;
    - its subset size, uniform on ``subset`` (inclusive);
    - its haplotypes, distinct and uniform;
    - its region length in bp, ``exp(U(log lo, log hi))``, turned into
      ``round(bp / bp_per_site)`` sites (at least 2, at most ``n_sites``);
    - its first site, uniform among the starts that fit.

    Returns ``(queries, arrival offsets in seconds)``.
    """
    rng = np.random.default_rng([seed, 3])
    qs, arr = [], []
    t = 0.0
    for _ in range(n_queries):
        t += rng.exponential(1.0 / rate)
        m = int(rng.integers(subset[0], subset[1] + 1))
        h = np.sort(rng.choice(n_haps, size=m, replace=False))
        bp = math.exp(rng.uniform(math.log(region_bp[0]), math.log(region_bp[1])))
        S = int(min(n_sites, max(2, round(bp / bp_per_site))))
        lo = int(rng.integers(0, n_sites - S + 1))
        qs.append(Query(h, lo, lo + S, core_block_sites, flank_sites))
        arr.append(t)
    return qs, np.array(arr)


def run_open_loop(server, queries, offsets, on_finish=None, sleep=20e-6):
    """Submit ``queries`` at ``start + offsets`` (host clock) while polling
    the server, until all have finished. ``on_finish(q)`` runs for each
    finished query (the default releases its memory). Returns the start
    time."""
    start = server.clock()
    k, n, done = 0, len(queries), 0
    while done < n:
        now = server.clock()
        while k < n and start + offsets[k] <= now:
            server.submit(queries[k], arrival=start + offsets[k])
            k += 1
        got = server.poll()
        for q in got:
            (on_finish or server.release)(q)
        done += len(got)
        if not got:
            wait = sleep if k >= n else min(sleep, start + offsets[k] - server.clock())
            if wait > 0:
                time.sleep(wait)
    return start
