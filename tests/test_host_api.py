"""Host-side API tests (no GPU): argument contract, validation, auto-sizing.

Restates the reference's fake-backend tests (``tests/unit/test_infer.py``
:98-144, :323-593, :666-675) against ``gamma_smc_cu_amd``: the native
``_core.FlowContext`` and ``_core.cuda_mem_info`` are replaced by fakes, as
the reference's tests do, so the Python layer is checked without a device.
"""

import importlib
import math
import re

import numpy as np
import pytest

import gamma_smc_cu_amd
infer_mod = importlib.import_module("gamma_smc_cu_amd.infer")
from gamma_smc_cu_amd.infer import _estimate_scaled_params, _recommend_core_block_size
from oracle import oracle as O


@pytest.fixture(scope="module")
def genotype_data():
    from gamma_smc_cu_amd import synthetic

    return synthetic.cohort(20, 217, 100_000, 43)


class _Recorder:
    def __init__(self, calls, S):
        self.calls = calls
        self.S = S

    def make(self):
        calls, S = self.calls, self.S

        class FakeFlowContext:
            def __init__(self, G, positions, Ne, mu, rho, flow_field_path, cache_steps):
                calls["init"] = dict(shape=G.shape, positions=len(positions), Ne=Ne, mu=mu, rho=rho,
                                     flow_field_path=flow_field_path, cache_steps=cache_steps)

            def run_fb(self, pairs, mean_only=True, return_posterior=False):
                calls["run_fb"] = dict(pairs=list(pairs), mean_only=mean_only, return_posterior=return_posterior)
                return {"mean": np.ones((S, len(pairs)), np.float32)}

            def run_fb_blockwise(self, pairs, core_block_sites, flank_sites, pair_batch_size, max_streams,
                                 mean_only, return_posterior=False):
                calls["run"] = dict(pairs=list(pairs), core_block_sites=core_block_sites, flank_sites=flank_sites,
                                    pair_batch_size=pair_batch_size, max_streams=max_streams, mean_only=mean_only,
                                    return_posterior=return_posterior)
                return {"mean": np.ones((S, len(pairs)), np.float32),
                        "blocks": np.array([[0, S, 0, S]], np.int32)}

        return FakeFlowContext


class TestAutoEstimateTheta:
    """Known answers of ``tests/unit/test_infer.py:98-144``."""

    def test_uses_4ne_convention_for_rho(self):
        G = np.array([[0, 0, 1, 1], [0, 1, 1, 1], [0, 0, 0, 1], [1, 0, 0, 1]], np.uint8)
        pos = np.array([10.0, 20.0, 30.0, 40.0])
        Ne, mu, rho = 10_000.0, 2.0e-8, 5.0e-8
        for fn in (_estimate_scaled_params, O.estimate_scaled_params):
            em, er = fn(G, pos, mu, rho, Ne)
            pi_hat = 1.0 / (pos[-1] + 1.0)
            np.testing.assert_allclose(4.0 * Ne * em, pi_hat)
            np.testing.assert_allclose(4.0 * Ne * er, pi_hat * rho / mu)

    def test_counts_leading_span_like_gamma_smc(self):
        G = np.array([[0, 1], [1, 1], [0, 0], [0, 1]], np.uint8)
        pos = np.array([100.0, 200.0])
        Ne, mu, rho = 10_000.0, 2.0e-8, 5.0e-8
        for fn in (_estimate_scaled_params, O.estimate_scaled_params):
            em, er = fn(G, pos, mu, rho, Ne)
            np.testing.assert_allclose(4.0 * Ne * em, 1.0 / 201.0)
            np.testing.assert_allclose(4.0 * Ne * er, (1.0 / 201.0) * rho / mu)

    def test_degenerate_inputs_keep_rates(self):
        G = np.array([[0, 1, 1]], np.uint8)
        assert _estimate_scaled_params(G, np.array([1.0, 2.0, 3.0]), 1e-8, 2e-8, 1e4) == (1e-8, 2e-8)
        G3 = np.zeros((3, 2), np.uint8)
        assert _estimate_scaled_params(G3, np.array([1.0, 2.0]), 1e-8, 2e-8, 1e4) == (1e-8, 2e-8)
        G4 = np.zeros((4, 2), np.uint8)
        assert _estimate_scaled_params(G4, np.array([1.0, 2.0]), 1e-8, 2e-8, 1e4) == (1e-8, 2e-8)

    def test_matches_oracle_on_random_cohort(self, genotype_data):
        G, pos = genotype_data
        assert _estimate_scaled_params(G, pos, 1.25e-8, 1e-8, 1e4) == O.estimate_scaled_params(
            G, pos, 1.25e-8, 1e-8, 1e4)


class TestHostLogic:
    def test_filter_segregating_matches_oracle(self, genotype_data):
        G, pos = genotype_data
        Gp = np.concatenate([np.zeros((20, 1), np.uint8), G, np.ones((20, 1), np.uint8)], axis=1)
        pp = np.concatenate([[pos[0] - 1], pos, [pos[-1] + 1]])
        g1, p1, k1 = infer_mod._filter_segregating(Gp, pp)
        g2, p2, k2 = O.filter_segregating(Gp, pp)
        np.testing.assert_array_equal(g1, G)
        np.testing.assert_array_equal(p1, pos)
        np.testing.assert_array_equal(k1, k2)
        g3, p3, _ = infer_mod._filter_segregating(G, pos)
        assert g3 is G and p3 is pos  # no copy when every site segregates

    def test_default_pair_order(self):
        assert O.all_pairs(4) == [(1, 0), (2, 0), (2, 1), (3, 0), (3, 1), (3, 2)]

    def test_pair_index_roundtrip(self):
        # tests/unit/test_pair_indexing.py:29-50 and include/gamma_smc_cu/types.h:56-68
        for n in (5, 10, 50):
            p = 0
            for i in range(1, n):
                for j in range(i):
                    assert O.ij_to_pair(i, j) == p == O.ij_to_pair(j, i)
                    assert O.pair_to_ij(p) == (i, j)
                    assert O.all_pairs(n)[p] == (i, j)
                    p += 1
            assert O.ij_to_pair(n - 1, n - 2) == n * (n - 1) // 2 - 1

    @pytest.mark.parametrize("S,core,flank", [(100, 30, 8), (217, 217, 0), (1, 5, 3), (402853, 32768, 2048),
                                              (10, 3, 0), (0, 4, 1)])
    def test_block_windows_closed_form(self, S, core, flank):
        got = gamma_smc_cu_amd._core.block_windows(S, core, flank)
        np.testing.assert_array_equal(got, O.block_windows(S, core, flank))
        if S:
            assert got[0, 0] == 0 and got[-1, 1] == S
            np.testing.assert_array_equal(got[1:, 0], got[:-1, 1])
            np.testing.assert_array_equal(got[:, 2], np.maximum(0, got[:, 0] - flank))
            np.testing.assert_array_equal(got[:, 3], np.minimum(S, got[:, 1] + flank))
        assert got.shape[0] == math.ceil(S / core)

    def test_block_windows_rejects_bad_args(self):
        with pytest.raises(ValueError, match="core_block_sites"):
            gamma_smc_cu_amd._core.block_windows(10, 0, 1)
        with pytest.raises(ValueError, match="flank_sites"):
            gamma_smc_cu_amd._core.block_windows(10, 2, -1)

    def test_recommend_core_block_size(self):
        # 217 sites + 32 GB -> one block (test_infer.py:499-500)
        assert _recommend_core_block_size(217, 2, True, 64, 32 * 10**9)[0] == 217
        # chromosome-scale on a 288 GB device -> 32768-site cores
        assert _recommend_core_block_size(402_853, 63_190, True, 2048, 287 * 10**9)[0] == 32768
        assert _recommend_core_block_size(5000, 10, True, 2048, 1024)[0] == 1024
        for args in [(402_853, 63_190, True, 2048, 40 * 10**9), (50_000, 499_500, False, 2048, 287 * 10**9),
                     (3000, 0, True, 64, 10**9)]:
            assert _recommend_core_block_size(*args) == O.recommend_core_block_size(*args)


class TestInferBlockwiseContract:
    """The reference's fake-backend tests (test_infer.py:323-593)."""

    def test_requires_explicit_pairs(self, genotype_data):
        G, pos = genotype_data
        with pytest.raises(ValueError, match="explicit pairs"):
            gamma_smc_cu_amd.infer_blockwise(G, pos, flow_field_path="dummy-flow-field.txt")

    def test_rejects_invalid_pair_batch_size(self, genotype_data):
        G, pos = genotype_data
        with pytest.raises(ValueError, match="pair_batch_size"):
            gamma_smc_cu_amd.infer_blockwise(G, pos, pairs=[(0, 1)], flow_field_path="dummy", pair_batch_size=0)

    def test_rejects_invalid_max_streams(self, genotype_data):
        G, pos = genotype_data
        with pytest.raises(ValueError, match="max_streams"):
            gamma_smc_cu_amd.infer_blockwise(G, pos, pairs=[(0, 1)], flow_field_path="dummy", max_streams=0)

    def test_rejects_negative_flank(self, genotype_data):
        G, pos = genotype_data
        with pytest.raises(ValueError, match="flank_sites"):
            gamma_smc_cu_amd.infer_blockwise(G, pos, pairs=[(0, 1)], flow_field_path="dummy", flank_sites=-1)

    def test_passes_blockwise_args(self, monkeypatch, genotype_data):
        G, pos = genotype_data
        pairs = [(0, 1), (2, 3)]
        calls = {}
        monkeypatch.setattr(gamma_smc_cu_amd._core, "FlowContext", _Recorder(calls, len(pos)).make())
        result = gamma_smc_cu_amd.infer_blockwise(
            G, pos, pairs=pairs, flow_field_path="dummy-flow-field.txt", core_block_sites=256, flank_sites=64,
            pair_batch_size=32, max_streams=3, mean_only=False)
        assert result["mean"].shape == (len(pos), len(pairs))
        assert result["pairs"] == pairs
        np.testing.assert_array_equal(result["positions"], pos)
        np.testing.assert_array_equal(result["blocks"], [[0, len(pos), 0, len(pos)]])
        assert calls["init"]["shape"] == G.shape
        assert calls["init"]["cache_steps"] == 0
        assert calls["run"] == {"pairs": pairs, "core_block_sites": 256, "flank_sites": 64, "pair_batch_size": 32,
                                "max_streams": 3, "mean_only": False, "return_posterior": False}

    def test_rejects_zero_flank_with_multi_block(self, genotype_data):
        G, pos = genotype_data
        with pytest.raises(ValueError, match="flank_sites=0"):
            gamma_smc_cu_amd.infer_blockwise(G, pos, pairs=[(0, 1)], flow_field_path="dummy",
                                             core_block_sites=10, flank_sites=0)

    def test_rejects_non_int_core_block_sites(self, genotype_data):
        G, pos = genotype_data
        with pytest.raises(TypeError, match="core_block_sites"):
            gamma_smc_cu_amd.infer_blockwise(G, pos, pairs=[(0, 1)], flow_field_path="dummy",
                                             core_block_sites="not-a-number")

    def test_rejects_posterior_with_streams(self, genotype_data):
        G, pos = genotype_data
        with pytest.raises(ValueError, match="return_posterior"):
            gamma_smc_cu_amd.infer_blockwise(G, pos, pairs=[(0, 1)], flow_field_path="dummy",
                                             core_block_sites=G.shape[1], flank_sites=0, return_posterior=True,
                                             max_streams=2)

    def test_auto_core_block_sites_uses_recommendation(self, monkeypatch, genotype_data):
        G, pos = genotype_data
        calls = {}
        monkeypatch.setattr(gamma_smc_cu_amd._core, "cuda_mem_info", lambda: (32 * 10**9, 40 * 10**9))
        monkeypatch.setattr(gamma_smc_cu_amd._core, "FlowContext", _Recorder(calls, len(pos)).make())
        gamma_smc_cu_amd.infer_blockwise(G, pos, pairs=[(0, 1), (2, 3)], flow_field_path="dummy",
                                         core_block_sites="auto", flank_sites=64)
        assert calls["run"]["core_block_sites"] == G.shape[1]
        assert calls["run"]["flank_sites"] == 64

    def test_warns_when_block_exceeds_memory(self, monkeypatch, genotype_data):
        G, pos = genotype_data
        monkeypatch.setattr(gamma_smc_cu_amd._core, "cuda_mem_info", lambda: (1024, 1024))
        monkeypatch.setattr(gamma_smc_cu_amd._core, "FlowContext", _Recorder({}, len(pos)).make())
        with pytest.warns(UserWarning, match="GPU memory"):
            gamma_smc_cu_amd.infer_blockwise(G, pos, pairs=[(0, 1)], flow_field_path="dummy",
                                             core_block_sites=8192, flank_sites=2048)

    def test_auto_pair_batch_size_capped_by_n_pairs(self, monkeypatch, genotype_data):
        G, pos = genotype_data
        calls = {}
        monkeypatch.setattr(gamma_smc_cu_amd._core, "cuda_mem_info", lambda: (32 * 10**9, 40 * 10**9))
        monkeypatch.setattr(gamma_smc_cu_amd._core, "FlowContext", _Recorder(calls, len(pos)).make())
        gamma_smc_cu_amd.infer_blockwise(G, pos, pairs=[(0, 1), (2, 3), (4, 5)], flow_field_path="dummy",
                                         pair_batch_size=-1)
        assert calls["run"]["pair_batch_size"] == 3

    def test_auto_falls_back_when_query_fails(self, monkeypatch, genotype_data):
        G, pos = genotype_data

        def boom():
            raise RuntimeError("no device")

        calls = {}
        monkeypatch.setattr(gamma_smc_cu_amd._core, "cuda_mem_info", boom)
        monkeypatch.setattr(gamma_smc_cu_amd._core, "FlowContext", _Recorder(calls, len(pos)).make())
        with pytest.warns(UserWarning, match="could not query GPU"):
            gamma_smc_cu_amd.infer_blockwise(G, pos, pairs=[(0, 1)], flow_field_path="dummy",
                                             core_block_sites="auto")
        assert calls["run"]["core_block_sites"] == 8192

    def test_auto_theta_reaches_the_context(self, monkeypatch, genotype_data):
        G, pos = genotype_data
        calls = {}
        monkeypatch.setattr(gamma_smc_cu_amd._core, "FlowContext", _Recorder(calls, len(pos)).make())
        gamma_smc_cu_amd.infer(G, pos, pairs=[(1, 0)], flow_field_path="dummy")
        assert (calls["init"]["mu"], calls["init"]["rho"]) == _estimate_scaled_params(G, pos, 1.25e-8, 1e-8, 1e4)
        gamma_smc_cu_amd.infer(G, pos, pairs=[(1, 0)], flow_field_path="dummy", auto_estimate_theta=False, Ne=5000)
        assert (calls["init"]["mu"], calls["init"]["rho"], calls["init"]["Ne"]) == (1.25e-8, 1e-8, 5000.0)


class TestInferContract:
    def test_default_pairs_and_result_keys(self, monkeypatch, genotype_data):
        G, pos = genotype_data
        calls = {}
        monkeypatch.setattr(gamma_smc_cu_amd._core, "FlowContext", _Recorder(calls, len(pos)).make())
        r = gamma_smc_cu_amd.infer(G, pos, flow_field_path="dummy")
        assert calls["run_fb"]["pairs"] == O.all_pairs(20)
        assert r["pairs"] == O.all_pairs(20)
        assert r["tmrca"] is r["mean"]
        assert r["positions"].dtype == np.float64

    def test_rejects_mismatched_positions(self, genotype_data):
        G, pos = genotype_data
        with pytest.raises(ValueError, match="positions length"):
            gamma_smc_cu_amd.infer(G, pos[:-1], pairs=[(0, 1)])

    def test_requires_positions_for_matrix(self, genotype_data):
        G, _ = genotype_data
        with pytest.raises(ValueError, match="positions is required"):
            gamma_smc_cu_amd.infer(G)

    def test_tree_sequence_duck_typing(self, monkeypatch, genotype_data):
        G, pos = genotype_data

        class Var:
            def __init__(self, p):
                self.position = p

        class FakeTS:
            def genotype_matrix(self):
                return G.T.astype(np.int32)

            def variants(self):
                return (Var(p) for p in pos)

        calls = {}
        monkeypatch.setattr(gamma_smc_cu_amd._core, "FlowContext", _Recorder(calls, len(pos)).make())
        r = gamma_smc_cu_amd.infer(FakeTS(), pairs=[(0, 1)], flow_field_path="dummy")
        assert calls["init"]["shape"] == G.shape
        np.testing.assert_array_equal(r["positions"], pos)

    def test_pairs_normalized_to_int_tuples(self, monkeypatch, genotype_data):
        G, pos = genotype_data
        calls = {}
        monkeypatch.setattr(gamma_smc_cu_amd._core, "FlowContext", _Recorder(calls, len(pos)).make())
        r = gamma_smc_cu_amd.infer(G, pos, pairs=np.array([[3, 1], [0, 2]]), flow_field_path="dummy")
        assert r["pairs"] == [(3, 1), (0, 2)] and all(type(i) is int for p in r["pairs"] for i in p)


class TestCABI:
    def test_library_exports_every_declared_symbol(self):
        import ctypes
        import os

        from gamma_smc_cu_amd import _core

        header = open(os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include",
                                   "gamma_smc_amd.h")).read()
        declared = set(re.findall(r"^\w[\w\s\*]*?\b(gsmc_\w+)\(", header, flags=re.M))
        assert declared == set(_core.EXPORTS)
        lib = ctypes.CDLL(_core.LIB_PATH)
        for name in declared:
            assert hasattr(lib, name), name

    def test_errors_map_to_python_exceptions(self):
        from gamma_smc_cu_amd import _core

        with pytest.raises(RuntimeError, match="flow field"):
            _core.build_tables("/nonexistent/field.txt", 10, 1e4, 1e-8, 1e-8)
        with pytest.raises(ValueError, match="cache_steps"):
            _core.build_tables(infer_mod._resolve_flow_field_path(None), 0, 1e4, 1e-8, 1e-8)

    def test_unknown_flag_bits_are_refused_before_any_device_call(self):
        """Entry points taking run flags refuse bits outside GSMC_CI ..
        GSMC_Q4 (the library's internal launch flags included) first, so
        the call fails with GSMC_EINVAL and a message without touching a
        device."""
        from gamma_smc_cu_amd import _core

        L = _core.lib()
        for bit in (32, 256, 1 << 20):
            assert L.gsmc_run_fb(None, None, None, 0, bit, None, None, None, None, None) == -1
            assert b"unknown flag bits" in L.gsmc_last_error()
            assert L.gsmc_run_fb_device(None, None, None, 0, 8, 0, bit, None, None, None, None, None, None) == -1
            assert L.gsmc_batch_launch(None, None, 0, bit, None) == -1
            assert L.gsmc_run_fwd(None, None, None, 0, bit, None, None, None) == -1
            assert b"unknown flag bits" in L.gsmc_last_error()
