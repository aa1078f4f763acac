"""Two-stage demographic decoding (``python/gamma_smc_cu/demographic.py:247-286``).

Stage 1 decodes a few calibration pairs with the constant-Ne flow field and
turns their posterior means into a piecewise-constant N(t)
(:func:`flow_field.estimate_demography`). Stage 2 regenerates the flow
field for that N(t) in FP64 (:func:`flow_field.generate_flow_field` with
``Ne_values``) and builds the decoder context on it. ``run_fb``,
``run_fwd`` and ``run_fb_summary`` then decode under the inferred
demography.
"""

from __future__ import annotations

import os
import tempfile

import numpy as np

from . import _core, flow_field


class DemographicFlowContext:
    def __init__(self, G, positions, Ne_init, mu, rho, flow_field_path=flow_field.DEFAULT_PATH,
                 n_calibration_pairs=50, n_epochs=20, n_steps=500, cache_steps=0, gpu_id=0, verbose=True):
        G = np.ascontiguousarray(G, np.uint8)
        n = G.shape[0]
        self.S = len(positions)
        self.Ne_init = float(Ne_init)
        _core.set_device(gpu_id)
        # calibration pairs: distinct (a, b), a < b, drawn from default_rng(42)
        rng = np.random.default_rng(42)
        n_cal = min(int(n_calibration_pairs), n * (n - 1) // 2)
        chosen = set()
        while len(chosen) < n_cal:
            a, b = sorted(rng.choice(n, 2, replace=False))
            chosen.add((int(a), int(b)))
        if verbose:
            print("Stage 1: estimating demography...", flush=True)
        ctx = _core.FlowContext(G, positions, float(Ne_init), mu, rho, flow_field_path, cache_steps)
        means = ctx.run_fb(sorted(chosen), mean_only=True)["mean"]
        ctx.close()
        self.Ne_values, self.epoch_boundaries = flow_field.estimate_demography(means, Ne_init, n_epochs=n_epochs)
        if verbose:
            print("Stage 2: generating the demographic flow field...", flush=True)
        u, v = flow_field.generate_flow_field(n_steps=n_steps, Ne_values=self.Ne_values,
                                              epoch_boundaries=self.epoch_boundaries, Ne_ref=float(Ne_init))
        self._dir = tempfile.mkdtemp(prefix="gsmc_ff_")
        self.flow_field_path = os.path.join(self._dir, "demographic_flow_field.txt")
        flow_field.write_flow_field(self.flow_field_path, u, v)
        self.ctx = _core.FlowContext(G, positions, float(Ne_init), mu, rho, self.flow_field_path, cache_steps)

    def run_fb(self, pairs, mean_only=True, return_posterior=False):
        return self.ctx.run_fb(pairs, mean_only, return_posterior)

    def run_fwd(self, pairs, mean_only=True):
        return self.ctx.run_fwd(pairs, mean_only)

    def run_fb_summary(self, pairs):
        return self.ctx.run_fb_summary(pairs)

    @property
    def device_id(self):
        return self.ctx.device_id

    @property
    def demography(self):
        return self.Ne_values.copy(), self.epoch_boundaries.copy()

    def __del__(self):
        try:
            if os.path.exists(self.flow_field_path):
                os.unlink(self.flow_field_path)
            os.rmdir(self._dir)
        except Exception:
            pass
