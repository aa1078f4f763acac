"""Known-answer pins of the oracle that leave no tolerance.

* ``orc_log_codes`` against a closed form of ``gamma_smc.cu:23-39``
  (``q = (log(max(mean, 1)) - log_min) * inv_range * levels + 0.5``,
  truncated and clamped to ``[0, levels]``) evaluated in exact rational
  arithmetic, hand-computed codes, and the uint4 nibble packing of
  ``gamma_smc.cu:185-205`` (even site low nibble, odd last site alone).
* The walk's table choices (``gamma_smc_flow.cu:471-514`` segments,
  ``:567-609`` / ``:632-669`` forward, ``:678-839`` backward): synthetic
  tables whose values lie below the grid box, so every lookup from a state
  below the box reads node (0, 0) with zero weights and returns that table's
  code exactly. The forward state at each site then names the (layer, slab)
  the forward step used, and the posterior alpha / beta name the table of
  the last backward lookup. A wrong layer, a swapped hom / het slab, a
  swapped direction or a misplaced N-bp stretch changes a code, and the
  comparison is bit for bit.
"""

from fractions import Fraction

import numpy as np
import pytest

from oracle import oracle as O


# ---------------------------------------------------------------- codes


def _round_f32(x):
    """Round a Fraction to the nearest float32 (ties to even)."""
    if x == 0:
        return np.float32(0.0)
    sign = -1 if x < 0 else 1
    x = abs(x)
    e = 0
    while Fraction(2) ** e > x:
        e -= 1
    while Fraction(2) ** (e + 1) <= x:
        e += 1
    ulp = Fraction(2) ** (e - 23)
    q, r = divmod(x, ulp)
    if r > ulp / 2 or (r == ulp / 2 and q % 2 == 1):
        q += 1
    return np.float32(sign * float(q * ulp))


def _code_closed_form(mean, Ne, levels):
    """Code of one FP32 mean: the float32 natural log (libm's, correctly
    rounded for these arguments; checked against float64 below), scaled by
    the float32 1/log(20 Ne), then the fused scale-and-add rounded once."""
    m = np.float32(max(np.float32(mean), np.float32(1.0)))
    lg = np.float32(np.log(np.float64(m)))
    inv = np.float32(np.float32(1.0) / np.float32(np.log(np.float32(20.0) * np.float32(Ne))))
    x = np.float32(lg * inv)
    q = _round_f32(Fraction(float(x)) * int(levels) + Fraction(1, 2))
    q = min(max(float(q), 0.0), float(levels))
    return int(q)


@pytest.mark.parametrize("bits", [8, 4])
def test_log_codes_closed_form(bits):
    Ne = 1e4
    levels = 255 if bits == 8 else 15
    rng = np.random.default_rng(5)
    top = float(np.float32(20.0) * np.float32(Ne))
    means = np.concatenate([
        10 ** rng.uniform(-1, 6, 400),           # below 1 (code 0) to far above 20 Ne (clamped)
        [0.0, 0.5, 1.0, np.sqrt(top), top, 10 * top, np.float32(3.4e38)],
    ]).astype(np.float32)
    if bits == 8:
        got = O.log_codes(means.reshape(-1, 1), 8, Ne).reshape(-1)
        for mv, g in zip(means, got):
            assert g == _code_closed_form(mv, Ne, levels), (mv, g)
    else:
        # 4-bit codes of one column, one code per row before packing
        codes = np.array([_code_closed_form(mv, Ne, levels) for mv in means], np.uint8)
        for S in (len(means), len(means) - 1):  # even and odd row counts
            packed = O.log_codes(means[:S].reshape(-1, 1), 4, Ne)
            assert packed.shape == ((S + 1) // 2, 1)
            for t in range(packed.shape[0]):
                lo = int(codes[2 * t])
                hi = int(codes[2 * t + 1]) if 2 * t + 1 < S else 0
                assert int(packed[t, 0]) == (lo | (hi << 4)), (S, t)


def test_log_codes_hand_computed():
    Ne = 1e4
    top = np.float32(20.0) * np.float32(Ne)
    # log(max(m, 1)) = 0 -> code 0; m = 20 Ne -> x = 1 -> levels; beyond -> clamped
    means = np.array([[0.25, 1.0, top, 4 * top]], np.float32).T
    assert O.log_codes(means, 8, Ne).reshape(-1).tolist() == [0, 0, 255, 255]
    # half way in log space: x = 1/2 -> 127.5 + 0.5 = 128 (8 bit), 8.0 (4 bit)
    half = np.float32(np.sqrt(np.float64(top)))
    assert abs(float(np.log(np.float64(half)) / np.log(np.float64(top))) - 0.5) < 1e-7
    assert O.log_codes(np.array([[half]], np.float32), 8, Ne)[0, 0] in (127, 128)
    # two rows in one byte: row 0 low nibble, row 1 high nibble; odd last row alone
    q4 = O.log_codes(np.array([[1.0], [top], [top]], np.float32), 4, Ne)
    assert q4[:, 0].tolist() == [0 | (15 << 4), 15]


# ---------------------------------------------------------------- walk table choices

N_STEPS = 4


def _code_tables(N):
    """[10][N][2550] tables, every node of a layer equal to its code. Codes
    have at most 6 significant bits, so a lookup from the prior (0, 0) --
    whose cell weights have at most 18 fraction bits -- also returns them
    exactly. Stretch layers other than N - 1 are NaN: the walk must never
    read them."""
    T = np.full((10, N, 2550), np.nan, np.float32)
    T[0, N - 1] = -12.0   # cache[N-1] m: the N-bp stretch
    T[1, N - 1] = -7.0    # cache[N-1] c
    for L in range(N):
        for h in range(2):
            k = 2 * L + h
            T[2 + 2 * h, L] = -(8.0 + k / 4.0)   # fwd (hom, het) m
            T[3 + 2 * h, L] = -3.0               # fwd c
            T[6 + 2 * h, L] = -6.0               # bwd m
            T[7 + 2 * h, L] = -(3.0 + k / 4.0)   # bwd c
    return T


def _steps(delta):
    return max(1, int(np.rint(delta)))


def _fwd_code(seg, het, N):
    rem = seg - N * ((seg - 1) // N)  # the final table after (seg - 1) // N stretches
    k = 2 * (rem - 1) + het
    return np.float32(-(8.0 + k / 4.0)), np.float32(-3.0)


def _bwd_last_code(seg, het, N):
    """State after the backward segment of `seg` bp: the site's table first,
    then the stretches, so the last lookup is a stretch when seg > N."""
    if seg > N:
        return np.float32(-12.0), np.float32(-7.0)
    k = 2 * (seg - 1) + het
    return np.float32(-6.0), np.float32(-(3.0 + k / 4.0))


def _case():
    rng = np.random.default_rng(11)
    gaps = [0.0, 1, 2, 2.5, 3.5, 4, 5, 7, 8, 9, 11.49, 12, 13, 1, 3, 4.5, 6, 16, 1, 2]
    pos = 5.0 + np.cumsum(gaps)  # pos[0] = 5: the first site applies rint(6) = 6 > N bp
    S = len(pos)
    G = rng.integers(0, 2, size=(4, S)).astype(np.uint8)
    G[0, :] = 0
    G[1, ::3] = 1  # pair (1, 0): het at every third site
    return G, pos


def _expected(G, pos, pair, N, start, stop):
    het = ((G[pair[0]] ^ G[pair[1]]) & 1).astype(int)
    F, Y = [], []
    for s in range(start, stop):
        if s == 0:
            seg = _steps(pos[0] + 1.0)
        elif s == start:
            seg = 1  # block restart: one bp
        else:
            seg = _steps(pos[s] - pos[s - 1])
        F.append(_fwd_code(seg, het[s], N))
    for s in range(start, stop):
        if s == stop - 1:
            Y.append((np.float32(0.0), np.float32(0.0)))
        else:
            Y.append(_bwd_last_code(_steps(pos[s + 1] - pos[s]), het[s + 1], N))
    return F, Y


def _combine(f, y, Ne):
    p10 = lambda x: np.float32(O.fast_pow10(float(x)))  # noqa: E731 (pinned by test_oracle_pins)
    lb = np.float32(np.float32(-2.0) * y[1])
    ba, bb = p10(lb), p10(np.float32(lb - y[0]))
    lf = np.float32(np.float32(-2.0) * f[1])
    fa, fb = p10(lf), p10(np.float32(lf - f[0]))
    a = np.float32(np.float32(fa + ba) - np.float32(1.0))
    b = np.float32(np.float32(fb + bb) - np.float32(1.0))
    mean = np.float32(np.float32(a / max(b, np.float32(1e-10))) * np.float32(2.0 * Ne))
    return a, b, mean


def test_forward_states_name_their_tables():
    G, pos = _case()
    T = _code_tables(N_STEPS)
    for pair in [(1, 0), (2, 3), (3, 1)]:
        m, c = O.forward_states_f32(G, pos, pair, T)
        F, _ = _expected(G, pos, pair, N_STEPS, 0, G.shape[1])
        assert np.array_equal(m, np.array([f[0] for f in F], np.float32)), pair
        assert np.array_equal(c, np.array([f[1] for f in F], np.float32)), pair


@pytest.mark.parametrize("start,stop", [(0, None), (5, None), (3, 14)])
def test_posterior_names_forward_and_backward_tables(start, stop):
    G, pos = _case()
    S = G.shape[1]
    stop = S if stop is None else stop
    T = _code_tables(N_STEPS)
    Ne = 1e4
    pairs = [(1, 0), (2, 3), (3, 1), (2, 0)]
    out = O.fb_window_f32(G, pos, pairs, T, Ne, site_start=start, block_S=stop - start, posterior=True)
    for p, pair in enumerate(pairs):
        F, Y = _expected(G, pos, pair, N_STEPS, start, stop)
        for s in range(stop - start):
            a, b, mean = _combine(F[s], Y[s], Ne)
            assert out["posterior_alpha"][s, p] == a, (pair, start + s)
            assert out["posterior_beta"][s, p] == b, (pair, start + s)
            assert out["mean"][s, p] == mean, (pair, start + s)


def test_code_tables_catch_a_planted_layer_slip():
    """The KAT fails on a one-layer slip: the same comparison with the
    expected forward layer shifted by one at a single site disagrees."""
    G, pos = _case()
    T = _code_tables(N_STEPS)
    m, _ = O.forward_states_f32(G, pos, (1, 0), T)
    F, _ = _expected(G, pos, (1, 0), N_STEPS, 0, G.shape[1])
    want = np.array([f[0] for f in F], np.float32)
    s = 7
    seg = _steps(pos[s] - pos[s - 1])
    slipped = want.copy()
    slipped[s] = _fwd_code(seg + 1 if seg % N_STEPS else seg - 1, int((G[1, s] ^ G[0, s]) & 1), N_STEPS)[0]
    assert np.array_equal(m, want) and not np.array_equal(m, slipped)
