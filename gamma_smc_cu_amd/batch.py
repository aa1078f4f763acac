"""Batched decoding of ragged jobs in one launch.

A job is a set of pairs decoded over a site range ``[site_lo, site_hi)`` of
a resident cohort, exactly as ``infer_blockwise`` would decode the cohort
restricted to those sites (``core_block_sites`` / ``flank_sites`` windows;
``core_block_sites <= 0`` decodes the range as one sequence, like ``infer``).
:func:`decode_batch` turns every job's (64-pair tile x block window) work
items into one kernel launch -- a static grid or persistent waves draining a
device queue, per the context's tuning -- so small jobs share the device
instead of leaving it idle. :func:`decode_plain` runs the same jobs one
launch each, the comparison mode. Both give bit-identical outputs: lanes
never mix pairs, so a job's result does not depend on its batch.
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from . import _core
from .device import DeviceArray


@dataclass
class Job:
    site_lo: int
    site_hi: int
    pairs: object  # (P, 2) int array or list of (i, j)
    core_block_sites: int = 0
    flank_sites: int = 0

    def pair_array(self):
        a = np.asarray(self.pairs, dtype=np.int64).reshape(-1, 2)
        return np.ascontiguousarray(a, dtype=np.int32)

    @property
    def n_sites(self):
        return int(self.site_hi) - int(self.site_lo)


class Workspace:
    """A ``gsmc_batch``: buffers of one launch in flight."""

    def __init__(self, ctx):
        h = ctypes.c_void_p()
        _core.check(_core.lib().gsmc_batch_create(ctx.handle, ctypes.byref(h)))
        self.handle = h
        self.ctx = ctx

    def launch(self, jobs_struct, n_jobs, flags, stream=None):
        _core.check(_core.lib().gsmc_batch_launch(self.handle, jobs_struct, int(n_jobs), int(flags),
                                                  None if stream is None else stream.handle))

    def wait(self):
        _core.check(_core.lib().gsmc_batch_wait(self.handle))

    def reserve(self, checkpoint_bytes, max_items, max_jobs):
        """Size the buffers for the largest launch now; a reserved workspace
        keeps its checkpoint buffer at this size (``gsmc_batch_reserve``)."""
        _core.check(_core.lib().gsmc_batch_reserve(self.handle, int(checkpoint_bytes), int(max_items),
                                                   int(max_jobs)))

    def checkpoint_buffer_bytes(self):
        """Bytes the checkpoint / stored-state buffer holds now."""
        b = ctypes.c_uint64()
        _core.check(_core.lib().gsmc_batch_buffer_bytes(self.handle, ctypes.byref(b)))
        return int(b.value)

    def done(self):
        d = ctypes.c_int()
        _core.check(_core.lib().gsmc_batch_query(self.handle, ctypes.byref(d)))
        return bool(d.value)

    def close(self):
        if self.handle is not None:
            _core.lib().gsmc_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Graph:
    """A batched launch captured into a HIP graph (``gsmc_batch_capture``):
    planned and uploaded once, replayed with :meth:`launch` on a stream.
    Replays write the outputs the captured jobs point to; replays of one
    graph must not overlap."""

    def __init__(self, ctx, jobs_struct, n_jobs, flags=0):
        h = ctypes.c_void_p()
        _core.check(_core.lib().gsmc_batch_capture(ctx.handle, jobs_struct, int(n_jobs), int(flags), ctypes.byref(h)))
        self.handle = h
        self.ctx = ctx

    def launch(self, stream=None):
        _core.check(_core.lib().gsmc_graph_launch(self.handle, None if stream is None else stream.handle))

    def close(self):
        if self.handle is not None:
            _core.lib().gsmc_graph_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _validate(ctx, jobs):
    for k, j in enumerate(jobs):
        if not (0 <= j.site_lo <= j.site_hi <= ctx.n_sites):
            raise ValueError(f"job {k}: site range [{j.site_lo}, {j.site_hi}) outside [0, {ctx.n_sites})")
        pa = j.pair_array()
        if pa.size and (pa.min() < 0 or pa.max() >= ctx.n_haps):
            raise ValueError(f"job {k}: pair index out of range")
        if j.flank_sites < 0:
            raise ValueError(f"job {k}: flank_sites must be non-negative.")


class Staged:
    """Device pair ids and outputs of a job list, laid out back to back."""

    def __init__(self, jobs, mean_only=True, return_posterior=False):
        self.jobs = list(jobs)
        self.n_out = 1 + (0 if mean_only else 2) + (2 if return_posterior else 0)
        self.mean_only = mean_only
        self.return_posterior = return_posterior
        pa = [j.pair_array() for j in self.jobs]
        self.pair_off = np.cumsum([0] + [p.shape[0] for p in pa])
        self.out_off = np.cumsum([0] + [j.n_sites * p.shape[0] for j, p in zip(self.jobs, pa)])
        allp = np.concatenate(pa) if pa else np.zeros((0, 2), np.int32)
        self.d_pi = DeviceArray.from_host(np.ascontiguousarray(allp[:, 0]))
        self.d_pj = DeviceArray.from_host(np.ascontiguousarray(allp[:, 1]))
        self.total_out = int(self.out_off[-1])
        self.d_out = DeviceArray((self.n_out, max(self.total_out, 1)), np.float32)
        self.flags = (0 if mean_only else _core.GSMC_CI) | (_core.GSMC_POSTERIOR if return_posterior else 0)

    def structs(self, which=None):
        idx = range(len(self.jobs)) if which is None else which
        arr = (_core.Job * max(1, len(idx)))()
        plane = max(self.total_out, 1)
        for slot, k in enumerate(idx):
            j = self.jobs[k]
            s = arr[slot]
            s.site_lo, s.site_hi = int(j.site_lo), int(j.site_hi)
            s.core_block_sites, s.flank_sites = int(j.core_block_sites), int(j.flank_sites)
            s.pair_i = self.d_pi.offset(self.pair_off[k]).value
            s.pair_j = self.d_pj.offset(self.pair_off[k]).value
            s.n_pairs = int(self.pair_off[k + 1] - self.pair_off[k])
            outs = [self.d_out.offset(o * plane + self.out_off[k]).value for o in range(self.n_out)]
            s.mean = outs[0]
            if not self.mean_only:
                s.lower, s.upper = outs[1], outs[2]
            if self.return_posterior:
                s.alpha, s.beta = outs[-2], outs[-1]
        return arr, len(idx)

    def results(self):
        host = self.d_out.to_host()
        names = ["mean"] + ([] if self.mean_only else ["lower", "upper"]) + \
                (["posterior_alpha", "posterior_beta"] if self.return_posterior else [])
        res = []
        for k, j in enumerate(self.jobs):
            P = int(self.pair_off[k + 1] - self.pair_off[k])
            a, b = int(self.out_off[k]), int(self.out_off[k + 1])
            res.append({name: host[o, a:b].reshape(j.n_sites, P) for o, name in enumerate(names)})
        return res


def decode_batch(ctx, jobs, mean_only=True, return_posterior=False, workspace=None, stream=None):
    """Decode every job in one launch; returns one result dict per job."""
    _validate(ctx, jobs)
    st = Staged(jobs, mean_only, return_posterior)
    ws = workspace or Workspace(ctx)
    arr, n = st.structs()
    ws.launch(arr, n, st.flags, stream)
    ws.wait()
    return st.results()


def decode_plain(ctx, jobs, mean_only=True, return_posterior=False, workspace=None, stream=None):
    """Decode the jobs one launch each (the per-request mode)."""
    _validate(ctx, jobs)
    st = Staged(jobs, mean_only, return_posterior)
    ws = workspace or Workspace(ctx)
    for k in range(len(jobs)):
        arr, n = st.structs([k])
        ws.launch(arr, n, st.flags, stream)
        ws.wait()
    return st.results()
