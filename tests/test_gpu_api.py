"""GPU tests of the drop-in API, restating the reference's device-backed tests
(``tests/unit/test_infer.py:26-95,147-213,216-223,249-319,424-442,604-700``)
on synthetic cohorts (msprime is not installed), plus the committed golden
vectors (``tests/golden/golden_case.npz``), which the HIP path must
reproduce bit for bit.
"""

import os

import numpy as np
import pytest

import gamma_smc_cu_amd
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _case(msg):
    print(f"[case] {msg}", flush=True)


@pytest.fixture(scope="module")
def genotype_data():
    from gamma_smc_cu_amd import synthetic

    return synthetic.cohort(20, 480, 300_000, 42)


def _eq_bits(a, b):
    return a.shape == b.shape and np.array_equal(np.ascontiguousarray(a).view(np.uint32),
                                                 np.ascontiguousarray(b).view(np.uint32))


def test_golden_fixture_bit_exact():
    from tests.golden import golden_case

    case = golden_case.load()
    pairs = [tuple(map(int, p)) for p in case["pairs"]]
    ff = golden_case.FF
    _case("golden full")
    ctx = gamma_smc_cu_amd.FlowContext(case["G"], case["positions"], 1e4, float(case["mu"]), float(case["rho"]), ff, 0)
    r = ctx.run_fb(pairs, mean_only=False, return_posterior=True)
    for k in ("mean", "lower", "upper", "posterior_alpha", "posterior_beta"):
        assert _eq_bits(r[k], case[k]), k
    _case("golden blockwise")
    b = ctx.run_fb_blockwise(pairs, core_block_sites=golden_case.CORE, flank_sites=golden_case.FLANK)
    assert _eq_bits(b["mean"], case["blockwise_mean"])
    np.testing.assert_array_equal(b["blocks"], case["blocks"])


class TestInferShapes:
    def test_all_pairs_shape_and_order(self, genotype_data):
        G, pos = genotype_data
        _case("infer all pairs")
        r = gamma_smc_cu_amd.infer(G, pos)
        n, S = G.shape
        assert r["mean"].shape == (S, n * (n - 1) // 2)
        assert r["pairs"] == O.all_pairs(n)
        assert r["mean"].dtype == np.float32 and r["positions"].dtype == np.float64
        assert np.all(r["mean"] > 0)

    def test_specific_pairs_and_custom_params(self, genotype_data):
        G, pos = genotype_data
        _case("infer custom params")
        r = gamma_smc_cu_amd.infer(G, pos, mu=1e-8, rho=1e-8, Ne=5000, pairs=[(0, 1)])
        assert r["mean"].shape == (G.shape[1], 1) and np.all(r["mean"] > 0)
        assert r["pairs"] == [(0, 1)]

    def test_empty_pairs(self, genotype_data):
        G, pos = genotype_data
        assert gamma_smc_cu_amd.infer(G, pos, pairs=[])["mean"].shape == (G.shape[1], 0)


class TestModes:
    def test_mean_only(self, genotype_data):
        G, pos = genotype_data
        r = gamma_smc_cu_amd.infer(G, pos, pairs=[(0, 1), (2, 3)], mean_only=True)
        assert "lower" not in r and "posterior_alpha" not in r

    def test_ci_ordering(self, genotype_data):
        G, pos = genotype_data
        _case("infer ci")
        r = gamma_smc_cu_amd.infer(G, pos, pairs=[(0, 1), (2, 3)], mean_only=False)
        assert np.all(r["lower"] <= r["mean"]) and np.all(r["mean"] <= r["upper"])

    def test_posterior_reconstructs_mean(self, genotype_data):
        G, pos = genotype_data
        _case("infer posterior")
        r = gamma_smc_cu_amd.infer(G, pos, pairs=[(0, 1), (2, 3)], mean_only=False, return_posterior=True)
        recon = r["posterior_alpha"] / r["posterior_beta"] * 2.0 * 1e4
        np.testing.assert_allclose(recon, r["mean"], rtol=1e-4, atol=1e-2)
        base = gamma_smc_cu_amd.infer(G, pos, pairs=[(0, 1), (2, 3)])
        np.testing.assert_array_equal(base["mean"], r["mean"])


class TestSegregatingFilter:
    def _pad(self, G, pos):
        n = G.shape[0]
        Gp = np.concatenate([np.zeros((n, 1), np.uint8), G, np.ones((n, 1), np.uint8)], axis=1)
        pp = np.concatenate([[pos[0] - 1.0], pos, [pos[-1] + 1.0]])
        return Gp, pp

    def test_filter_matches_unpadded(self, genotype_data):
        G, pos = genotype_data
        Gp, pp = self._pad(G, pos)
        _case("filter infer")
        a = gamma_smc_cu_amd.infer(G, pos, pairs=[(0, 1), (2, 3)])
        b = gamma_smc_cu_amd.infer(Gp, pp, pairs=[(0, 1), (2, 3)])
        np.testing.assert_array_equal(a["mean"], b["mean"])
        np.testing.assert_array_equal(b["positions"], pos)

    def test_blockwise_filter_matches_unpadded(self, genotype_data):
        G, pos = genotype_data
        Gp, pp = self._pad(G, pos)
        _case("filter blockwise")
        a = gamma_smc_cu_amd.infer_blockwise(G, pos, pairs=[(0, 1), (2, 3)], core_block_sites=G.shape[1],
                                             flank_sites=0)
        b = gamma_smc_cu_amd.infer_blockwise(Gp, pp, pairs=[(0, 1), (2, 3)], core_block_sites=G.shape[1],
                                             flank_sites=0)
        np.testing.assert_array_equal(a["mean"], b["mean"])


class TestBlockwise:
    def test_single_block_matches_full_sequence(self, genotype_data):
        G, pos = genotype_data
        _case("blockwise single block")
        full = gamma_smc_cu_amd.infer(G, pos, pairs=[(0, 1), (2, 3)])
        blk = gamma_smc_cu_amd.infer_blockwise(G, pos, pairs=[(0, 1), (2, 3)], core_block_sites=G.shape[1],
                                               flank_sites=0, pair_batch_size=1)
        np.testing.assert_array_equal(blk["mean"], full["mean"])
        np.testing.assert_array_equal(blk["blocks"], [[0, G.shape[1], 0, G.shape[1]]])

    def test_posterior_and_ci_blockwise(self, genotype_data):
        G, pos = genotype_data
        _case("blockwise posterior + ci")
        r = gamma_smc_cu_amd.infer_blockwise(G, pos, pairs=[(0, 1), (2, 3)], core_block_sites=G.shape[1],
                                             flank_sites=0, mean_only=False, return_posterior=True)
        np.testing.assert_allclose(r["posterior_alpha"] / r["posterior_beta"] * 2e4, r["mean"], rtol=1e-4,
                                   atol=1e-2)
        assert np.all(r["lower"] <= r["mean"] + 1e-3) and np.all(r["mean"] <= r["upper"] + 1e-3)

    def test_streamed_matches_single_stream(self, genotype_data):
        G, pos = genotype_data
        kw = dict(pairs=[(0, 1), (2, 3), (4, 5)], core_block_sites=G.shape[1] // 2,
                  flank_sites=min(64, G.shape[1] // 4), pair_batch_size=2)
        _case("blockwise streams")
        a = gamma_smc_cu_amd.infer_blockwise(G, pos, max_streams=1, **kw)
        b = gamma_smc_cu_amd.infer_blockwise(G, pos, max_streams=2, **kw)
        np.testing.assert_array_equal(a["mean"], b["mean"])

    def test_auto_core_block_sites_on_device(self, genotype_data):
        G, pos = genotype_data
        _case("blockwise auto")
        r = gamma_smc_cu_amd.infer_blockwise(G, pos, pairs=[(0, 1)], verbose=True)
        assert r["blocks"].shape == (1, 4)  # a few hundred sites fit one block

    def test_flanked_blocks_match_oracle_and_converge(self, genotype_data, flow_field_path):
        G, pos = genotype_data
        pairs = O.all_pairs(G.shape[0])[:64]
        _case("blockwise flanked vs oracle")
        r = gamma_smc_cu_amd.infer_blockwise(G, pos, pairs=pairs, core_block_sites=120, flank_sites=100)
        mu, rho = O.estimate_scaled_params(G, pos, 1.25e-8, 1e-8, 1e4)
        T = O.tables_f32(flow_field_path, 1000, 1e4, mu, rho)
        ref = O.fb_blockwise_f32(G, pos, pairs, T, 1e4, 120, 100)
        assert _eq_bits(r["mean"], ref["mean"])
        full = gamma_smc_cu_amd.infer(G, pos, pairs=pairs)["mean"]
        assert np.median(np.abs(np.log(r["mean"]) - np.log(full))) < 1e-3


def test_fp64_tables_and_error_report(genotype_data, flow_field_path):
    """FP64-built tables stay within the stated tolerance of the FP32 path,
    and the FP32 GPU posterior's error against the FP64 oracle is bounded:
    |log mean_fp32 - log mean_fp64| <= 1e-2 everywhere, median <= 1e-3."""
    G, pos = genotype_data
    pairs = O.all_pairs(G.shape[0])[:40]
    mu, rho = O.estimate_scaled_params(G, pos, 1.25e-8, 1e-8, 1e4)
    _case("fp64 tables")
    c32 = gamma_smc_cu_amd.FlowContext(G, pos, 1e4, mu, rho, flow_field_path, 0)
    c64 = gamma_smc_cu_amd.FlowContext(G, pos, 1e4, mu, rho, flow_field_path, 0, table_precision="fp64")
    m32 = c32.run_fb(pairs)["mean"]
    m64t = c64.run_fb(pairs)["mean"]
    T64 = O.tables_f64(flow_field_path, 1000, 1e4, mu, rho)
    exact = O.fb_window_f64(G, pos, pairs, T64, 1e4)["mean"]
    for m in (m32, m64t):
        err = np.abs(np.log(m) - np.log(exact))
        assert err.max() <= 1e-2 and np.median(err) <= 1e-3


def test_demographic_two_stage_context(genotype_data, flow_field_path):
    """Two-stage demographic decoding: the stage-2 field is the FP64 field
    of the estimated N(t), and decoding on it is bit-exact to the oracle
    built from the same field file."""
    from gamma_smc_cu_amd import flow_field as F
    from gamma_smc_cu_amd.demographic import DemographicFlowContext

    G, pos = genotype_data
    mu, rho = O.estimate_scaled_params(G, pos, 1.25e-8, 1e-8, 1e4)
    _case("demographic context")
    d = DemographicFlowContext(G, pos, 1e4, mu, rho, flow_field_path, n_calibration_pairs=20, n_epochs=5,
                               n_steps=100, verbose=False)
    Ne, bounds = d.demography
    assert Ne.shape == (5,) and bounds.shape == (6,) and np.all(Ne >= 100)
    u, v, _ = F.read_flow_field(d.flow_field_path)
    gu, gv = F.generate_flow_field(n_steps=100, Ne_values=Ne, epoch_boundaries=bounds, Ne_ref=1e4)
    np.testing.assert_allclose(u, gu, atol=1e-10)
    np.testing.assert_allclose(v, gv, atol=1e-10)
    pairs = [(0, 1), (2, 3), (5, 4)]
    r = d.run_fb(pairs)
    T = O.tables_f32(d.flow_field_path, 1000, 1e4, mu, rho)
    assert _eq_bits(r["mean"], O.fb_window_f32(G, pos, pairs, T, 1e4)["mean"])
    s = d.run_fb_summary(pairs)
    assert s["n_pairs"] == 3 and np.isfinite(s["site_mean"]).all()
    f = d.run_fwd(pairs)
    assert f["mean"].shape == r["mean"].shape and (f["mean"] > 0).all()
