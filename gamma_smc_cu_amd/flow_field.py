"""Base (u, v) flow field for the Gamma-SMC transition, generated in FP64.

The decoder's recombination step moves a Gamma(alpha, beta) posterior, held
as (m, c) = (log10 mean, log10 CV) in scaled time T / (2 Ne), along a 2-D
vector field (u, v) tabulated on a 51 x 50 grid (m in [-5, 2], c in [-2, 0]).
The upstream package ships that field as ``default_flow_field.txt``; the
file is not distributed with the reference, so this module regenerates it
from the published construction (Schweiger & Durbin 2023) as restated by
``python/gamma_smc_cu/demographic.py:17-44,133-208`` (constant-Ne branch):

* For each grid node the change of the marginal TMRCA density under one
  SMC' recombination step, ``delta(t)``, is evaluated on a two-part time
  grid (``demographic.py:133-143``) with the closed form of
  ``_dist_diff_const_Ne`` (``demographic.py:17-44``).
* ``delta`` is projected, by weighted least squares, onto the derivatives of
  the Gamma density with respect to log10(alpha) and log10(beta)
  (``demographic.py:186-199``); the projection gives d(log10 alpha) and
  d(log10 beta), which map to (u, v) = (r0 - r1, -r0 / 2).
* The field is divided by the calibration constant 5.302
  (``demographic.py:14,206-207``).

Everything here is FP64 (numpy/scipy). The text format written by
:func:`write_flow_field` is the one ``load_flow_field`` parses
(``src/flow_field.cpp:8-54``; writer ``demographic.py:215-220``).
"""

from __future__ import annotations

import os

import numpy as np

MEAN_N = 51
CV_N = 50
CALIBRATION = 5.302

DEFAULT_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)),
                            "default_flow_field.txt")


def _time_grid(alpha: float, beta: float, n_steps: int, resolve: bool = False) -> np.ndarray:
    """Two-part time grid: dense over the Gamma bulk, then an Exp(1) tail.

    Follows ``demographic.py:133-143``: the first part spans
    [1e-10, Gamma quantile 0.999] (at least 0.1) in ``n_steps`` intervals,
    the second extends by -ln(1e-3) in ``n_steps`` further intervals.

    ``resolve=True`` is this module's fallback for nodes whose density is
    narrower than that grid's spacing (small mean and small CV): the first
    part then spans the Gamma's own [0.001, 0.999] quantile range.
    """
    from scipy.special import gammaincinv

    try:
        hi = float(gammaincinv(alpha, 1.0 - 1e-3)) / beta
        if not np.isfinite(hi):
            raise ValueError
    except Exception:
        hi = alpha / beta * 5.0
    if resolve:
        lo = float(gammaincinv(alpha, 1e-3)) / beta
        if not (np.isfinite(lo) and 0.0 < lo < hi):
            lo = 1e-10
        head = np.linspace(lo, hi, n_steps + 1)
    else:
        hi = max(hi, 0.1)
        head = np.linspace(1e-10, hi, n_steps + 1)
    tail_step = -np.log(1e-3) / max(n_steps - 1, 1)
    tail = head[-1] + tail_step * np.arange(1, n_steps + 1)
    return np.concatenate([head, tail])


def _density_change(t: np.ndarray, alpha: float, beta: float) -> np.ndarray:
    """Change of the TMRCA density at times ``t`` after one recombination step.

    Closed form for a constant population size (``demographic.py:17-44``):
    re-coalescence mass from below and above the breakpoint (incomplete-Gamma
    terms), minus the mass that leaves the current time, plus the Exp tail.
    """
    from scipy.special import gammainc, gammaincc, gammaln, hyp1f1

    t = np.maximum(np.asarray(t, dtype=np.float64), 1e-15)
    below = beta - 1.0
    above = beta + 1.0
    if below > 1e-10:
        log_ratio = alpha * (np.log(beta) - np.log(below))
        from_below = np.exp(log_ratio - t) * gammainc(alpha, below * t)
    else:
        # beta <= 1: the incomplete Gamma with a non-positive rate is
        # written through the confluent hypergeometric 1F1(a; a+1; x).
        log_base = alpha * np.log(t * beta) - t - gammaln(alpha + 1.0)
        h = hyp1f1(alpha, alpha + 1.0, -below * t)
        with np.errstate(divide="ignore", invalid="ignore"):
            log_h = np.where(h > 0, np.log(h), -np.inf)
        from_below = np.exp(log_base + log_h)
        from_below = np.where(np.isfinite(from_below), from_below, 0.0)
    from_above = np.exp(alpha * (np.log(beta) - np.log(above)) - t) \
        * gammainc(alpha, above * t)
    log_pdf = alpha * np.log(beta) - gammaln(alpha) + (alpha - 1.0) * np.log(t) - beta * t
    pdf = np.exp(np.where(log_pdf > -700.0, log_pdf, -700.0))
    sf = gammaincc(alpha, beta * t)
    e2 = np.exp(-2.0 * t)
    return (from_below - from_above) + (0.5 * e2 - 0.5 - t) * pdf + (1.0 - e2) * sf


def _epochs(Ne_values, epoch_boundaries, Ne_ref):
    """Piecewise-constant coalescence rate in scaled time s = T / (2 Ne_ref)
    (``demographic.py:149-158``): epoch k starts at boundary k (generations,
    the first at 0; the last epoch is open-ended), rate lam_k = Ne_ref / N_k,
    and the cumulative rate Lam(s) is L_k + lam_k (s - t_k) inside epoch k."""
    M = len(Ne_values)
    t_b = np.zeros(M + 1)
    t_b[1:M] = np.asarray(epoch_boundaries, np.float64)[1:M] / (2.0 * Ne_ref)
    t_ext = np.append(t_b[:M], 1e15)
    lam = float(Ne_ref) / np.asarray(Ne_values, np.float64)
    L = np.zeros(M + 1)
    for k in range(1, M):
        L[k] = L[k - 1] + lam[k - 1] * (t_b[k] - t_b[k - 1])
    return lam, t_ext, L


def _density_change_demographic(t, alpha, beta, lam, t_ext, L):
    """Change of the TMRCA density after one recombination step when the
    coalescence rate is piecewise constant (``demographic.py:47-130``):

        e^{-Lam(t)} (I+(t) - I-(t)) + (e^{-2 Lam}/2 - 1/2 - Lam) pdf + (1 - e^{-2 Lam}) sf

    with I+-(t) = int_0^t e^{+-Lam(s)} Gamma_pdf(s) ds summed epoch by epoch in
    closed form: inside epoch k the integrand is e^{+-(L_k - lam_k t_k)} times
    a Gamma density of rate c = beta -+ lam_k, so each piece is
    e^{...} (beta / c)^alpha [P(alpha, c b) - P(alpha, c a)]. As upstream:
    pieces whose scale is below e^-500 are dropped, c ~ 0 uses the power
    form, and pieces with c < 0 (a rate above beta) are left out."""
    from scipy.special import gammainc, gammaincc, gammaln

    t = np.maximum(np.asarray(t, dtype=np.float64), 1e-15)
    M = len(lam)
    k_of_t = np.clip(np.searchsorted(t_ext[:M], t, side="right") - 1, 0, M - 1)
    Lam = L[k_of_t] + lam[k_of_t] * (t - t_ext[k_of_t])
    ip = np.zeros_like(t)
    im = np.zeros_like(t)
    for k in range(M):
        a_k = t_ext[k]
        b_k = np.minimum(t, t_ext[k + 1] if k < M - 1 else 1e15)
        on = t > a_k + 1e-15
        if not on.any():
            break
        shift = L[k] - lam[k] * t_ext[k]
        c = beta - lam[k]
        if c > 1e-10:
            log_scale = shift + alpha * (np.log(beta) - np.log(c))
            if log_scale > -500.0:
                lo = gammainc(alpha, c * max(a_k, 1e-30)) if a_k > 1e-15 else 0.0
                hi = np.where(on, gammainc(alpha, c * np.maximum(b_k, 1e-30)), 0.0)
                ip += np.where(on, np.exp(log_scale) * (hi - lo), 0.0)
        elif abs(c) <= 1e-10 and alpha > 0:
            coef = np.exp(shift + alpha * np.log(beta) - gammaln(alpha))
            hi = np.where(on, np.power(np.maximum(b_k, 1e-30), alpha), 0.0)
            lo = max(a_k, 1e-30) ** alpha if a_k > 1e-15 else 0.0
            ip += np.where(on, coef * (hi - lo) / alpha, 0.0)
        c = beta + lam[k]
        log_scale = -shift + alpha * (np.log(beta) - np.log(c))
        if log_scale > -500.0:
            lo = gammainc(alpha, c * max(a_k, 1e-30)) if a_k > 1e-15 else 0.0
            hi = np.where(on, gammainc(alpha, c * np.maximum(b_k, 1e-30)), 0.0)
            im += np.where(on, np.exp(log_scale) * (hi - lo), 0.0)
    log_pdf = alpha * np.log(beta) - gammaln(alpha) + (alpha - 1.0) * np.log(t) - beta * t
    pdf = np.exp(np.where(log_pdf > -700.0, log_pdf, -700.0))
    sf = gammaincc(alpha, beta * t)
    e2 = np.exp(-2.0 * Lam)
    return np.exp(-Lam) * (ip - im) + (0.5 * e2 - 0.5 - Lam) * pdf + (1.0 - e2) * sf


def generate_flow_field(mean_min=1e-5, mean_max=1e2, cv_min=1e-2, cv_max=1.0,
                        mean_n=MEAN_N, cv_n=CV_N, n_steps=500, Ne_values=None, epoch_boundaries=None,
                        Ne_ref=10000.0):
    """Return the (u, v) field, each ``(mean_n, cv_n)`` float64, mean-major.

    Grid nodes are log-spaced: ``log10(mean)`` over [log10 mean_min,
    log10 mean_max] and ``log10(cv)`` over [log10 cv_min, log10 cv_max]
    (``demographic.py:161-162``). Nodes with a degenerate Gamma (alpha or
    beta below 1e-8, alpha above 1e8) or fewer than three usable time points
    keep a zero flow, as upstream does (``demographic.py:174-195``).

    With ``Ne_values`` (one population size per epoch) and
    ``epoch_boundaries`` (epoch start times in generations, the first 0),
    the density change is the piecewise-constant-rate form of
    :func:`_density_change_demographic`, in time scaled by ``2 Ne_ref``
    (``demographic.py:146-208``).

    Departure from upstream, on purpose: where upstream's time grid leaves
    the density unresolved -- fewer than 8 grid points within a factor 1e3
    of the density's peak, or a log-density below -672 everywhere, so the
    fit sees only the ``exp(-700)`` floor and returns values up to 1e293 --
    the node is refitted on the quantile-resolved grid of
    :func:`_time_grid`. Every other node is computed exactly as upstream.
    """
    from scipy.special import digamma, gammaln

    demo = None
    if Ne_values is not None:
        if epoch_boundaries is None or len(epoch_boundaries) < len(Ne_values):
            raise ValueError("epoch_boundaries needs one start time per epoch")
        demo = _epochs(Ne_values, epoch_boundaries, Ne_ref)
    lm = np.linspace(np.log10(mean_min), np.log10(mean_max), mean_n)
    lc = np.linspace(np.log10(cv_min), np.log10(cv_max), cv_n)
    ln10 = np.log(10.0)
    u = np.zeros((mean_n, cv_n))
    v = np.zeros((mean_n, cv_n))
    for i in range(mean_n):
        mean = 10.0 ** lm[i]
        for j in range(cv_n):
            cv = 10.0 ** lc[j]
            sd = cv * mean
            alpha = (mean / sd) ** 2
            beta = mean / sd ** 2
            if alpha < 1e-8 or beta < 1e-8 or alpha > 1e8:
                continue
            for resolve in (False, True):
                t = _time_grid(alpha, beta, n_steps, resolve)
                tt = np.maximum(t, 1e-30)
                log_pdf = alpha * np.log(beta) - gammaln(alpha) + (alpha - 1.0) * np.log(tt) - beta * t
                peak = log_pdf.max()
                if peak > -700.0 - np.log(1e-12) and np.count_nonzero(log_pdf > peak - np.log(1e3)) >= 8:
                    break
            w = np.sqrt(np.append(np.diff(t), 0.0))
            with np.errstate(over="ignore", invalid="ignore"):
                change = _density_change(t, alpha, beta) if demo is None \
                    else _density_change_demographic(t, alpha, beta, *demo)
            rhs = change * w
            pdf = np.exp(np.where(log_pdf > -700.0, log_pdf, -700.0))
            # d pdf / d log10(alpha) and d pdf / d log10(beta), weighted.
            d_la = pdf * (np.log(beta) - digamma(alpha) + np.log(tt)) * w * alpha / ln10
            d_lb = pdf * (alpha / beta - t) * w * beta / ln10
            keep = pdf > pdf.max() * 1e-12
            if keep.sum() < 3:
                continue
            design = np.column_stack([d_la, d_lb])[keep]
            if not (np.isfinite(rhs[keep]).all() and np.isfinite(design).all()):
                continue  # an overflowing integral: upstream's fit raises and keeps a zero flow
            try:
                sol = np.linalg.lstsq(design, rhs[keep], rcond=None)[0]
            except np.linalg.LinAlgError:
                continue
            # d(log10 mean) = d(log10 a) - d(log10 b); d(log10 cv) = -d(log10 a)/2
            u[i, j] = sol[0] - sol[1]
            v[i, j] = -0.5 * sol[0]
    return u / CALIBRATION, v / CALIBRATION


def estimate_demography(tmrca_means, Ne_init, n_epochs=20, t_max=None):
    """Stage 1 of the two-stage demographic fit (``demographic.py:227-245``):
    epochs with quadratically spaced boundaries on [0, t_max] generations
    (default 10 Ne_init); in each epoch holding at least 10 posterior means,
    N = max(100, survival / (2 density)) from the empirical density and
    survival of the means; other epochs keep Ne_init. Returns
    ``(Ne_values, boundaries)`` (``n_epochs`` sizes, ``n_epochs + 1``
    boundaries)."""
    if t_max is None:
        t_max = 10.0 * Ne_init
    vals = np.asarray(tmrca_means, np.float64).ravel()
    vals = vals[np.isfinite(vals) & (vals > 0)]
    bounds = t_max * np.linspace(0.0, 1.0, n_epochs + 1) ** 2
    bounds[0] = 0.0
    Ne = np.full(n_epochs, float(Ne_init))
    for k in range(n_epochs):
        lo, hi = bounds[k], bounds[k + 1]
        count = np.count_nonzero((vals >= lo) & (vals < hi))
        if count < 10 or hi <= lo:
            continue
        density = count / (vals.size * (hi - lo))
        survival = np.mean(vals >= lo)
        if survival > 0.01:
            Ne[k] = max(100.0, 1.0 / (2.0 * density / survival))
    return Ne, bounds


def write_flow_field(path, u, v, mean_min=1e-5, mean_max=1e2, cv_min=1e-2, cv_max=1.0):
    """Write the text format parsed by ``load_flow_field``.

    Two header lines ``<min> <max> <n>`` (linear grid bounds, Python ``repr``
    of the floats as upstream writes them), then the u rows and the v rows,
    one grid row per line, ``%.10f`` per value (``demographic.py:215-220``).
    """
    mn, cn = u.shape
    with open(path, "w") as fh:
        fh.write(f"{mean_min} {mean_max} {mn}\n{cv_min} {cv_max} {cn}\n")
        for row in u:
            fh.write(" ".join(f"{x:.10f}" for x in row) + "\n")
        for row in v:
            fh.write(" ".join(f"{x:.10f}" for x in row) + "\n")


def read_flow_field(path):
    """Parse a flow-field file into ``(u, v, header)`` (float64 arrays).

    ``header`` is ``((mean_min, mean_max, mean_n), (cv_min, cv_max, cv_n))``
    in linear units. Raises ``ValueError`` when the grid is not 51 x 50, the
    only shape the decoder accepts (``src/flow_field.cpp:17-33``).
    """
    with open(path) as fh:
        parts = fh.read().split()
    if len(parts) < 6:
        raise ValueError(f"flow field {path!r}: truncated header")
    mm, mx, mn = float(parts[0]), float(parts[1]), int(parts[2])
    cm, cx, cn = float(parts[3]), float(parts[4]), int(parts[5])
    if mn != MEAN_N or cn != CV_N:
        raise ValueError(f"flow field {path!r}: grid {mn}x{cn}, expected {MEAN_N}x{CV_N}")
    body = np.array([float(x) for x in parts[6:6 + 2 * mn * cn]])
    if body.size != 2 * mn * cn:
        raise ValueError(f"flow field {path!r}: expected {2 * mn * cn} values, got {body.size}")
    return body[:mn * cn].reshape(mn, cn), body[mn * cn:].reshape(mn, cn), ((mm, mx, mn), (cm, cx, cn))


if __name__ == "__main__":  # regenerate the packaged default field
    import sys

    out = sys.argv[1] if len(sys.argv) > 1 else DEFAULT_PATH
    uu, vv = generate_flow_field()
    write_flow_field(out, uu, vv)
    print(f"wrote {out}")
