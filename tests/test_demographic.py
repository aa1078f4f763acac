"""Demographic flow fields (host, FP64) and the two-stage context.

No reference output exists for the demographic field (the reference package
cannot be imported here), so the generator is pinned by what the
mathematics fixes: one epoch at Ne_ref reproduces the constant-Ne closed
form; splitting an epoch into equal-size epochs changes nothing; the
epoch bookkeeping integrates the rate exactly; the N(t) estimator recovers
a constant size from exponential coalescence times.
"""

import numpy as np
import pytest

from gamma_smc_cu_amd import flow_field as F

GRID = dict(mean_n=8, cv_n=6)


def _beta_grid():
    lm = np.linspace(-5, 2, GRID["mean_n"])
    lc = np.linspace(-2, 0, GRID["cv_n"])
    m, c = np.meshgrid(10.0 ** lm, 10.0 ** lc, indexing="ij")
    return m / (c * m) ** 2


def test_one_epoch_equals_constant_ne():
    u0, v0 = F.generate_flow_field(**GRID)
    u1, v1 = F.generate_flow_field(**GRID, Ne_values=[1e4], epoch_boundaries=[0.0], Ne_ref=1e4)
    ok = _beta_grid() > 1.01  # below, both forms follow upstream's own special cases
    np.testing.assert_allclose(u1[ok], u0[ok], rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(v1[ok], v0[ok], rtol=1e-9, atol=1e-12)


def test_equal_epochs_are_one_epoch():
    u1, v1 = F.generate_flow_field(**GRID, Ne_values=[1e4], epoch_boundaries=[0.0], Ne_ref=1e4)
    u3, v3 = F.generate_flow_field(**GRID, Ne_values=[1e4] * 3, epoch_boundaries=[0.0, 3e3, 2e4], Ne_ref=1e4)
    np.testing.assert_allclose(u3, u1, rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(v3, v1, rtol=1e-9, atol=1e-12)


def test_epoch_rate_integral():
    lam, t_ext, L = F._epochs([5e3, 2e4, 1e4], [0.0, 4e3, 1e4], 1e4)
    np.testing.assert_allclose(lam, [2.0, 0.5, 1.0])
    np.testing.assert_allclose(t_ext[:3], [0.0, 0.2, 0.5])
    # Lam at the epoch starts: 0, 2 * 0.2, 0.4 + 0.5 * 0.3
    np.testing.assert_allclose(L[:3], [0.0, 0.4, 0.55])


def test_size_change_moves_the_field():
    u1, _ = F.generate_flow_field(**GRID, Ne_values=[1e4], epoch_boundaries=[0.0], Ne_ref=1e4)
    u2, _ = F.generate_flow_field(**GRID, Ne_values=[2e3, 1e4, 5e4], epoch_boundaries=[0.0, 3e3, 2e4], Ne_ref=1e4)
    assert np.isfinite(u2).all() and np.abs(u2 - u1).max() > 1e-3


def test_estimate_demography_recovers_constant_size():
    rng = np.random.default_rng(3)
    Ne = 8000.0
    t = rng.exponential(2.0 * Ne, size=200_000)  # pairwise TMRCA, constant N
    est, bounds = F.estimate_demography(t, Ne, n_epochs=10)
    assert bounds[0] == 0.0 and bounds[-1] == pytest.approx(10 * Ne) and len(est) == 10
    np.testing.assert_allclose(bounds, 10 * Ne * np.linspace(0, 1, 11) ** 2)
    # the estimator is a per-epoch hazard: for exponential times its
    # expectation is dt / (2 (1 - exp(-dt / 2N))) (N for a narrow epoch)
    dt = np.diff(bounds)
    expect = dt / (2.0 * (1.0 - np.exp(-dt / (2.0 * Ne))))
    np.testing.assert_allclose(est[:8], expect[:8], rtol=0.03)
    assert abs(est[0] - Ne) < 0.05 * Ne
    # too few samples: the initial size is kept
    est2, _ = F.estimate_demography(t[:5], 1234.0, n_epochs=4)
    np.testing.assert_array_equal(est2, 1234.0)
