"""Launch-shape autotuner for the decode kernel.

Every candidate computes bit-identical results (the launch shape decides
only which wave decodes which pair tile of which block window, and how
many sites a checkpoint window spans), so the tuner is free to pick the
fastest. Candidates sweep:

* ``checkpoint_k`` -- sites per checkpoint window, which is also the length
  of the forward window recomputed into registers (8 or 16), or 1: every
  forward state stored by the forward walk and read back by the backward
  walk, one lookup per site fewer for 16 bytes of HBM traffic per
  pair-site (two kernels per phase of items with the static schedule, one
  fused kernel with the persistent one);
* ``pairs_per_lane`` -- 1, 2 or 4 pairs per lane, i.e. 64, 128 or 256
  pairs per 64-lane workgroup (more pairs give every phase more independent
  lookup chains per lane, at the cost of registers and, at 4, of the LDS
  that holds the recomputed forward states; 4 is built for checkpoint_k 8);
* ``schedule`` -- a static grid of work items, ordered per XCD or not,
  persistent waves draining a device queue, or a dynamic grid (one
  workgroup per item, each taking the next item of the queue).

The core block size is not a tuning knob: block windows change the
posterior (each window restarts from the prior), so they stay what the
caller asked for.
"""

from __future__ import annotations

import itertools
import time

CANDIDATES = [
    dict(checkpoint_k=k, pairs_per_lane=np_, schedule=sch, xcd_order=xcd)
    for k, np_, (sch, xcd) in itertools.product(
        (8, 16), (1, 2), (("static", False), ("static", True), ("persistent", False), ("dynamic", False)))
] + [dict(checkpoint_k=8, pairs_per_lane=4, schedule=sch, xcd_order=False) for sch in ("static", "persistent")] \
  + [dict(checkpoint_k=1, pairs_per_lane=np_, schedule=sch, xcd_order=False)
     for np_ in (1, 2) for sch in ("static", "dynamic")] \
  + [dict(checkpoint_k=1, pairs_per_lane=1, schedule="persistent", xcd_order=False)]


def tune(ctx, run_once, candidates=None, repeats=1, log=None):
    """Time ``run_once()`` (which launches one decode on ``ctx`` and returns
    its duration in ms) under each candidate; leave ``ctx`` set to the
    fastest and return ``(best, [(candidate, ms), ...])``."""
    results = []
    for cand in candidates or CANDIDATES:
        ctx.set_tuning(**cand)
        run_once()  # first launch of a shape uploads metadata and grows buffers
        ms = min(run_once() for _ in range(max(1, repeats)))
        results.append((cand, ms))
        if log:
            log(f"tune {cand} -> {ms:.2f} ms")
    best = min(results, key=lambda r: r[1])[0]
    ctx.set_tuning(**best)
    # the sweep leaves the workspaces sized for its largest candidate (the
    # stored-state shape holds a state slot per resident wave)
    ctx.release_buffers()
    return best, results


def timed(fn):
    t0 = time.perf_counter()
    fn()
    return (time.perf_counter() - t0) * 1e3
