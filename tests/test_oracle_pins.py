"""Pins of the CPU oracle against the reference and against mathematics.

* Tables: the oracle's FP32 restatement and the product's host builder are
  compared bit for bit with the reference's own ``build_flow_field_cache``
  compiled from ``src/flow_field.cpp`` (oracle/_ref), and with committed
  SHA-256 digests of the reference build's output.
* Entropy clip: the bisected CV threshold is checked against scipy's Gamma
  entropy (an independent formula).
* Device bilinear: exact on linear tables, keeps the upper boundary cell
  (the reference KAT ``tests/unit/test_gamma_smc_flow.py:140-158`` restated
  on the 51 x 50 grid).
* Fast 10^x: bit pattern against an exact-rational fused multiply-add and the
  construction's error bound.
* Walk: short walks recomputed independently with scipy's
  RegularGridInterpolator on the same tables, including gaps longer than
  the table depth and a block start.
* Posterior identities and FP64 agreement on random cohorts.
"""

import hashlib
import json
import os
from fractions import Fraction

import numpy as np
import pytest
from scipy.interpolate import RegularGridInterpolator

from oracle import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
FF = os.path.join(os.path.dirname(HERE), "gamma_smc_cu_amd", "default_flow_field.txt")
PARAMS = [(1000, 1e4, 1.25e-8, 1e-8), (37, 5000.0, 2e-8, 5e-8)]


def _digest(a):
    return hashlib.sha256(np.ascontiguousarray(a, np.float32).tobytes()).hexdigest()


@pytest.mark.parametrize("N,Ne,mu,rho", PARAMS)
def test_oracle_tables_match_reference_build(N, Ne, mu, rho):
    if not O.ref_available():
        pytest.skip("oracle/_ref not built (reference tree absent)")
    a = O.tables_f32(FF, N, Ne, mu, rho)
    r = O.ref_tables(FF, N, Ne, mu, rho)
    for k, name in enumerate(O.TABLE_NAMES):
        assert np.array_equal(a[k].view(np.uint32), r[k].view(np.uint32)), name


@pytest.mark.parametrize("N,Ne,mu,rho", PARAMS)
def test_tables_match_committed_reference_digests(N, Ne, mu, rho):
    with open(os.path.join(HERE, "golden", "ref_table_digests.json")) as fh:
        golden = json.load(fh)
    key = f"N={N},Ne={Ne},mu={mu},rho={rho}"
    a = O.tables_f32(FF, N, Ne, mu, rho)
    got = {name: _digest(a[k]) for k, name in enumerate(O.TABLE_NAMES)}
    assert got == golden[key]


@pytest.mark.parametrize("N,Ne,mu,rho", PARAMS)
def test_product_host_tables_match_oracle(N, Ne, mu, rho):
    from gamma_smc_cu_amd import _core

    fwd, bwd, stretch = _core.build_tables(FF, N, Ne, mu, rho)
    a = O.tables_f32(FF, N, Ne, mu, rho)
    pairs = {"fwd_hom_m": fwd[:, 0, :, 0], "fwd_hom_c": fwd[:, 0, :, 1], "fwd_het_m": fwd[:, 1, :, 0],
             "fwd_het_c": fwd[:, 1, :, 1], "bwd_hom_m": bwd[:, 0, :, 0], "bwd_hom_c": bwd[:, 0, :, 1],
             "bwd_het_m": bwd[:, 1, :, 0], "bwd_het_c": bwd[:, 1, :, 1]}
    for name, got in pairs.items():
        assert np.array_equal(got.view(np.uint32), a[O.TABLE_NAMES.index(name)].view(np.uint32)), name
    assert np.array_equal(stretch[:, 0], a[0][-1]) and np.array_equal(stretch[:, 1], a[1][-1])


def test_fp64_tables_close_to_fp32():
    from gamma_smc_cu_amd import _core

    f32 = _core.build_tables(FF, 200, 1e4, 1.25e-8, 1e-8, "fp32")
    f64 = _core.build_tables(FF, 200, 1e4, 1.25e-8, 1e-8, "fp64")
    for a, b in zip(f32, f64):
        assert np.abs(a - b).max() < 1e-3
    t64 = O.tables_f64(FF, 200, 1e4, 1.25e-8, 1e-8)
    assert np.abs(t64[2] - f64[0][:, 0, :, 0]).max() < 1e-6  # product FP64 vs oracle FP64 (fwd hom m)


def test_entropy_threshold_against_scipy():
    from scipy.stats import gamma

    tf, td = O.entropy_table()
    for i in (0, 400, 1250, 2500, 3750, 5000):
        m = i * 0.0008
        c = td[i]
        if -2.0 < c < 0.0:
            a = 10.0 ** (-2 * c)
            b = 10.0 ** (-2 * c - m)
            assert gamma(a, scale=1.0 / b).entropy() == pytest.approx(1.0, abs=1e-6)
    assert np.all(np.diff(td) <= 1e-12)  # threshold falls as the mean grows
    assert np.array_equal(tf, td.astype(np.float32))


def _grid_tables(f):
    t = np.zeros((51, 50), np.float32)
    for r in range(51):
        for q in range(50):
            t[r, q] = f(r, q)
    return t.ravel()


def test_device_bilinear_exact_on_linear_tables():
    tm = _grid_tables(lambda r, q: 3.0 * r + 7.0 * q)
    tc = _grid_tables(lambda r, q: -2.0 * r + 0.5 * q + 1.0)
    rng = np.random.default_rng(0)
    for m, c in zip(rng.uniform(-5, 2, 200), rng.uniform(-2, 0, 200)):
        om, oc = O.dev_bilinear(tm, tc, m, c)
        fm = (m + 5) / 0.14
        fc = (c + 2) / (2 / 49)
        assert om == pytest.approx(3 * fm + 7 * fc, abs=2e-4)
        assert oc == pytest.approx(-2 * fm + 0.5 * fc + 1, abs=2e-4)
    # clamped outside the box
    assert O.dev_bilinear(tm, tc, -9.0, -7.0) == (0.0, 1.0)


def test_device_bilinear_keeps_upper_boundary_cell():
    # reference KAT (tests/unit/test_gamma_smc_flow.py:140-158) on the
    # 51 x 50 grid: at the top edge the weight stays on the last cell.
    t = _grid_tables(lambda r, q: 10.0 * r + q)
    assert O.dev_bilinear(t, t, 2.0, 0.0) == (549.0, 549.0)
    assert O.dev_bilinear(t, t, 2.0, -2.0) == (500.0, 500.0)
    assert O.dev_bilinear(t, t, -5.0, 0.0) == (49.0, 49.0)


def _fma_f32(a, b, c):
    """Correctly rounded float32 fma via exact rationals."""
    x = Fraction(float(a)) * Fraction(float(b)) + Fraction(float(c))
    if x == 0:
        return np.float32(0.0)
    sign = -1 if x < 0 else 1
    x = abs(x)
    e = x.numerator.bit_length() - x.denominator.bit_length()
    while Fraction(2) ** e > x:
        e -= 1
    while Fraction(2) ** (e + 1) <= x:
        e += 1
    ulp = Fraction(2) ** (e - 23)
    q, r = divmod(x, ulp)
    if r > ulp / 2 or (r == ulp / 2 and q % 2 == 1):
        q += 1
    return np.float32(sign * float(q * ulp))


def test_fast_pow10_bits_and_error_bound():
    LN10, C1, C2 = np.float32(2.30258509299), np.float32(1064872507.1541044), np.float32(12102203.161561485)
    rng = np.random.default_rng(1)
    xs = np.concatenate([rng.uniform(-2, 9, 300), [0.0, 1.0, -1.0, 4.0]]).astype(np.float32)
    for x in xs:
        t = np.float32(x * LN10)
        y = _fma_f32(C2, t, C1)
        bits = np.int32(np.trunc(np.float64(y)))
        expect = np.array([bits], np.int32).view(np.float32)[0]
        assert O.fast_pow10(float(x)) == expect
        assert abs(expect / 10.0 ** float(x) - 1.0) < 0.07  # Schraudolph construction: a few % at most


def test_seg_steps_rounding():
    cases = {0.5: 1, 1.5: 2, 2.5: 2, 3.5: 4, 126.4: 126, 126.6: 127, -5.0: 1, 0.0: 1, 1e12: 2147483647}
    for d, want in cases.items():
        assert O.seg_steps(d) == want
    assert O.seg_steps(float("nan")) == 1


# ---------------------------------------------------------------- walk

MG = np.linspace(-5.0, 2.0, 51)
CG = np.linspace(-2.0, 0.0, 50)


def _interp(layer, m, c):
    """Independent bilinear: scipy on the (m, c) grid, float64."""
    m = min(max(m, -5.0), 2.0)
    c = min(max(c, -2.0), 0.0)
    om = RegularGridInterpolator((MG, CG), layer[:, 0].reshape(51, 50).astype(np.float64))([m, c])[0]
    oc = RegularGridInterpolator((MG, CG), layer[:, 1].reshape(51, 50).astype(np.float64))([m, c])[0]
    return om, oc


def _steps(delta, N):
    n = max(1, int(np.rint(delta)))
    q = (n - 1) // N
    return q, n - q * N


def _walk(G, pos, pair, fwd, bwd, stretch, N, Ne, start=0, stop=None):
    """Recursion A for one pair over [start, stop), as the reference kernels walk it
    (gamma_smc_flow.cu:567-609 forward, :678-754 backward, :632-669 /
    :762-839 block starts), with scipy interpolation."""
    stop = G.shape[1] if stop is None else stop
    i, j = pair
    het = (G[i] ^ G[j]) & 1
    F = []
    m = c = 0.0
    for s in range(start, stop):
        if s == 0:
            d = pos[0] + 1.0
        elif s == start:
            d = 1.0
        else:
            d = pos[s] - pos[s - 1]
        q, r = _steps(d, N)
        for _ in range(q):
            m, c = _interp(stretch, m, c)
        m, c = _interp(fwd[r - 1, het[s]], m, c)
        F.append((m, c))
    out = np.zeros(stop - start)
    ym = yc = 0.0
    for s in range(stop - 1, start - 1, -1):
        fm, fc = F[s - start]
        a = O.fast_pow10(-2 * fc) + O.fast_pow10(-2 * yc) - 1.0
        b = O.fast_pow10(-2 * fc - fm) + O.fast_pow10(-2 * yc - ym) - 1.0
        out[s - start] = a / max(b, 1e-10) * 2 * Ne
        if s == 0:
            d = pos[0] + 1.0
        elif s == start:
            d = 1.0
        else:
            d = pos[s] - pos[s - 1]
        q, r = _steps(d, N)
        ym, yc = _interp(bwd[r - 1, het[s]], ym, yc)
        for _ in range(q):
            ym, yc = _interp(stretch, ym, yc)
    return out


@pytest.mark.parametrize("N", [1000, 7])
def test_walk_against_independent_interpolation(N):
    from gamma_smc_cu_amd import _core

    Ne, mu, rho = 1e4, 1.25e-8, 1e-8
    fwd, bwd, stretch = _core.build_tables(FF, N, Ne, mu, rho)
    T = O.tables_f32(FF, N, Ne, mu, rho)
    G = np.array([[0, 1, 0, 1, 1, 0], [1, 1, 0, 0, 1, 1], [0, 0, 0, 1, 0, 1]], np.uint8)
    pos = np.array([499.0, 3499.0, 3500.0, 3620.4, 5001.5, 5003.0])
    for pair in [(1, 0), (2, 0), (2, 1), (1, 1)]:
        want = _walk(G, pos, pair, fwd, bwd, stretch, N, Ne)
        got = O.fb_window_f32(G, pos, [pair], T, Ne)["mean"][:, 0]
        np.testing.assert_allclose(got, want, rtol=2e-4)
        # a block window starting at site 2 (first step of one bp)
        want_b = _walk(G, pos, pair, fwd, bwd, stretch, N, Ne, start=2, stop=5)
        got_b = O.fb_window_f32(G, pos, [pair], T, Ne, site_start=2, block_S=3)["mean"][:, 0]
        np.testing.assert_allclose(got_b, want_b, rtol=2e-4)


@pytest.fixture(scope="module")
def walk_case():
    from gamma_smc_cu_amd import synthetic

    G, pos = synthetic.cohort(10, 900, 500_000, 4)
    mu, rho = O.estimate_scaled_params(G, pos, 1.25e-8, 1e-8, 1e4)
    T = O.tables_f32(FF, 1000, 1e4, mu, rho)
    return G, pos, mu, rho, T


def test_posterior_identities(walk_case):
    G, pos, mu, rho, T = walk_case
    pairs = O.all_pairs(G.shape[0]) + [(3, 3)]
    r = O.fb_window_f32(G, pos, pairs, T, 1e4, ci=True, posterior=True)
    np.testing.assert_allclose(r["posterior_alpha"] / r["posterior_beta"] * 2e4, r["mean"], rtol=1e-5)
    assert (r["lower"] <= r["mean"]).all() and (r["mean"] <= r["upper"]).all()
    assert (r["mean"] > 0).all() and np.isfinite(r["mean"]).all()
    # pair order does not matter: (i, j) and (j, i) see the same XOR
    s = O.fb_window_f32(G, pos, [(p[1], p[0]) for p in pairs], T, 1e4)
    assert np.array_equal(s["mean"], r["mean"])


def test_fp32_against_fp64_oracle(walk_case):
    G, pos, mu, rho, T = walk_case
    pairs = O.all_pairs(G.shape[0])
    r32 = O.fb_window_f32(G, pos, pairs, T, 1e4)["mean"]
    T64 = O.tables_f64(FF, 1000, 1e4, mu, rho)
    r64 = O.fb_window_f64(G, pos, pairs, T64, 1e4)["mean"]
    err = np.abs(np.log(r32) - np.log(r64))
    assert err.max() < 1e-2 and np.median(err) < 1e-3


def test_blockwise_single_block_equals_full(walk_case):
    G, pos, mu, rho, T = walk_case
    pairs = O.all_pairs(G.shape[0])[:20]
    full = O.fb_window_f32(G, pos, pairs, T, 1e4)
    blk = O.fb_blockwise_f32(G, pos, pairs, T, 1e4, G.shape[1], 0)
    assert np.array_equal(full["mean"], blk["mean"])
    np.testing.assert_array_equal(blk["blocks"], [[0, G.shape[1], 0, G.shape[1]]])


def test_blockwise_flanks_converge_to_full(walk_case):
    G, pos, mu, rho, T = walk_case
    pairs = O.all_pairs(G.shape[0])[:20]
    full = O.fb_window_f32(G, pos, pairs, T, 1e4)["mean"]
    blk = O.fb_blockwise_f32(G, pos, pairs, T, 1e4, 200, 256)["mean"]
    err = np.abs(np.log(blk) - np.log(full))
    assert np.median(err) < 1e-3
    no_flank = O.fb_blockwise_f32(G, pos, pairs, T, 1e4, 200, 0)["mean"]
    assert np.median(np.abs(np.log(no_flank) - np.log(full))) > np.median(err)


def test_golden_fixture_reproduced_by_oracle():
    from tests.golden import golden_case

    case = golden_case.load()
    got = golden_case.compute_with_oracle(case)
    for k, v in got.items():
        assert np.array_equal(v.view(np.uint32), case[k].view(np.uint32)), k


def test_site_summary_restatement():
    """The summary oracle is the plain per-site mean (to FP64 rounding of a
    reordered sum) and the exact min / max; a dropped group or lane would
    move the mean by far more than the tolerance."""
    rng = np.random.default_rng(4)
    for P in (1, 63, 64, 65, 200):
        m = np.exp(rng.normal(8.0, 1.5, size=(37, P))).astype(np.float32)
        r = O.site_summary(m)
        ref = m.astype(np.float64).mean(axis=1)
        np.testing.assert_allclose(r["site_mean"], ref.astype(np.float32), rtol=2e-7)
        np.testing.assert_array_equal(r["site_min"], m.min(axis=1))
        np.testing.assert_array_equal(r["site_max"], m.max(axis=1))
    # one group by hand: four in-order parts of 16, joined pairwise
    x = (np.arange(64, dtype=np.float64) * 1e-3 + 1.0).astype(np.float32)[None, :]
    v = x.astype(np.float64)[0]
    parts = []
    for p in range(4):
        acc = 0.0
        for i in range(16):
            acc = acc + v[16 * p + i]
        parts.append(acc)
    assert O.site_summary(x)["site_mean"][0] == np.float32(((parts[0] + parts[1]) + (parts[2] + parts[3])) / 64)


def test_forward_only_oracle_against_independent_walk():
    """The forward-only oracle (run_fwd's recursion) against a float64
    walk with scipy interpolation on the same FP16-rounded tables: gap
    (int)(delta + 0.5) with nothing at site 0 or for gaps rounding to 0,
    full N-bp layers first, explicit het emission, mean = 10^m * 2Ne."""
    Ne, mu, rho, N = 1e4, 1.25e-8, 1e-8, 1000
    T = O.tables_f32(FF, N, Ne, mu, rho)
    tm, tc = O.h2_tables(T)
    # the FP16 tables are the reference's cache[k] rounded to nearest even
    assert np.array_equal(tm, T[0].astype(np.float16).astype(np.float32))
    G = np.array([[0, 1, 0, 1, 1, 0, 1], [1, 1, 0, 0, 1, 1, 0], [0, 0, 0, 1, 0, 1, 1]], np.uint8)
    pos = np.array([499.0, 3499.0, 3499.2, 3620.5, 5001.5, 6001.5, 6002.0])
    layers = np.stack([tm, tc], axis=-1)  # (N, 2550, 2)
    for pair in [(1, 0), (2, 0), (2, 1), (1, 1)]:
        het = (G[pair[0]] ^ G[pair[1]]) & 1
        m = c = 0.0
        want = []
        for s in range(G.shape[1]):
            if s > 0:
                gap = int(pos[s] - pos[s - 1] + 0.5)
                if gap > 0:
                    while gap > N:
                        m, c = _interp(layers[N - 1], m, c)
                        gap -= N
                    m, c = _interp(layers[gap - 1], m, c)
            if het[s]:
                a_log = np.log10(10.0 ** (-2.0 * c) + 1.0)
                m, c = a_log - (-2.0 * c - m), -0.5 * a_log
            want.append(10.0 ** m * 2.0 * Ne)
        got = O.fwd_only_f32(G, pos, [pair], (tm, tc), Ne)["mean"][:, 0]
        np.testing.assert_allclose(got, want, rtol=2e-4)
