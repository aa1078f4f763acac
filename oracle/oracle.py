"""CPU oracle for the Gamma-SMC decoder path.

TEST INFRASTRUCTURE ONLY: imported by ``tests/``, by
``__graft_entry__.smoke()`` and by the ``cpu_baseline`` leg of ``bench.py``,
always as the checker, never as the thing measured or shipped. The product
package (``gamma_smc_cu_amd``) does not import it and fails loudly when its
HIP library is missing.

Contents:

* ctypes bindings to ``liboracle.so`` (``gsmc_oracle.c``): the FP32 and FP64
  restatements of the reference's table builder (``src/flow_field.cpp``) and
  cached walk (``src/kernels/gamma_smc_flow.cu``, recursion A);
* numpy restatements of the reference's host logic in
  ``python/gamma_smc_cu/infer.py`` (input coercion, segregating filter,
  auto-theta, default pairs, block sizing) and ``src/bindings.cpp``
  (block windows, blockwise stitching);
* numpy restatement of the per-gene accumulators of the reference's
  production caller (``analysis/genome_wide/infer_chromosome.py:157-227``);
* ``ref_tables`` -- the reference's own ``build_flow_field_cache`` compiled
  from ``/root/reference/src/flow_field.cpp`` into ``oracle/_ref`` (only
  where that tree exists), used to pin ``tables_f32``.

Pinning status per function is recorded in the tests that use it
(``tests/test_oracle_*.py``).
"""

from __future__ import annotations

import ctypes
import math
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_LIB_PATH = os.path.join(HERE, "_ref", "libgsmc_ref.so")
GRID = 51 * 50
N_TABLES = 10
TABLE_NAMES = ("cache_m", "cache_c", "fwd_hom_m", "fwd_hom_c", "fwd_het_m",
               "fwd_het_c", "bwd_hom_m", "bwd_hom_c", "bwd_het_m", "bwd_het_c")

_lib = None
_ref = None

_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")


def build(force=False):
    """Compile liboracle.so (and oracle/_ref when the reference exists)."""
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-s", "-C", HERE, "liboracle.so"], check=True)
    if os.path.exists("/root/reference/src/flow_field.cpp") and (force or not os.path.exists(REF_LIB_PATH)):
        subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        L.orc_load_flow_field.argtypes = [ctypes.c_char_p, _f32p, _f32p, _f32p]
        L.orc_build_tables_f32.argtypes = [_f32p, _f32p, _f32p, ctypes.c_int, ctypes.c_float, ctypes.c_float, _f32p]
        L.orc_build_tables_f64.argtypes = [_f32p, _f32p, _f32p, ctypes.c_int, ctypes.c_double, ctypes.c_double, _f64p]
        opt_f32 = ctypes.c_void_p
        L.orc_fb_window_f32.argtypes = [_u8p, ctypes.c_int64, ctypes.c_int64, _f64p, _i32p, _i32p, ctypes.c_int64,
                                        _f32p, ctypes.c_int, ctypes.c_float, ctypes.c_int64, ctypes.c_int64,
                                        _f32p, opt_f32, opt_f32, opt_f32, opt_f32]
        L.orc_fb_window_f64.argtypes = [_u8p, ctypes.c_int64, ctypes.c_int64, _f64p, _i32p, _i32p, ctypes.c_int64,
                                        _f64p, ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_int64,
                                        ctypes.c_int64, _f64p, ctypes.c_void_p, ctypes.c_void_p]
        L.orc_forward_states_f32.argtypes = [_u8p, ctypes.c_int64, ctypes.c_int64, _f64p, ctypes.c_int32,
                                             ctypes.c_int32, _f32p, ctypes.c_int, _f32p, _f32p]
        L.orc_fwd_only_f32.argtypes = [_u8p, ctypes.c_int64, ctypes.c_int64, _f64p, _i32p, _i32p, ctypes.c_int64,
                                       _f32p, _f32p, ctypes.c_int, ctypes.c_float, _f32p, ctypes.c_void_p,
                                       ctypes.c_void_p]
        L.orc_fwd_only_f64.argtypes = [_u8p, ctypes.c_int64, ctypes.c_int64, _f64p, _i32p, _i32p, ctypes.c_int64,
                                       _f64p, _f64p, ctypes.c_int, ctypes.c_double, _f64p]
        L.orc_log_codes.argtypes = [_f32p, ctypes.c_int64, ctypes.c_float, ctypes.c_float, _u8p]
        L.orc_log_codes.restype = None
        L.orc_fast_pow10.argtypes = [ctypes.c_float]
        L.orc_fast_pow10.restype = ctypes.c_float
        L.orc_seg_steps.argtypes = [ctypes.c_double]
        L.orc_seg_steps.restype = ctypes.c_int
        L.orc_dev_bilinear.argtypes = [_f32p, _f32p, ctypes.c_float, ctypes.c_float,
                                       ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float)]
        L.orc_entropy_table.argtypes = [_f32p, _f64p]
        _lib = L
    return _lib


def ref_available():
    return os.path.exists(REF_LIB_PATH)


def ref_lib():
    global _ref
    if _ref is None:
        L = ctypes.CDLL(REF_LIB_PATH)
        L.ref_build_tables.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_float, ctypes.c_float, _f32p, _f32p]
        L.ref_load_flow_field.argtypes = [ctypes.c_char_p, _f32p, _f32p, _f32p]
        _ref = L
    return _ref


# ---------------------------------------------------------------- tables

def load_flow_field(path):
    """(u, v, bounds) as the reference parses them (``flow_field.cpp:8-54``)."""
    u = np.zeros(GRID, np.float32)
    v = np.zeros(GRID, np.float32)
    b = np.zeros(4, np.float32)
    rc = lib().orc_load_flow_field(os.fsencode(path), u, v, b)
    if rc != 0:
        raise ValueError(f"oracle: cannot parse flow field {path!r} (code {rc})")
    return u, v, b


def scaled_rates(Ne, mu, rho):
    """``4.0f * Ne * rho`` and ``4.0f * Ne * mu`` in FP32 with float Ne, mu, rho
    (``bindings.cpp:2575,2044-2045``)."""
    ne = np.float32(Ne)
    return np.float32(np.float32(4.0) * ne * np.float32(rho)), np.float32(np.float32(4.0) * ne * np.float32(mu))


def tables_f32(path, n_steps, Ne, mu, rho):
    """FP32 tables, (10, n_steps, 2550) float32 in ``TABLE_NAMES`` order."""
    u, v, b = load_flow_field(path)
    srho, smu = scaled_rates(Ne, mu, rho)
    out = np.zeros(N_TABLES * n_steps * GRID, np.float32)
    lib().orc_build_tables_f32(u, v, b, int(n_steps), float(srho), float(smu), out)
    return out.reshape(N_TABLES, n_steps, GRID)


def tables_f64(path, n_steps, Ne, mu, rho):
    """FP64 tables: the same construction in double (FP64 scaled rates)."""
    u, v, b = load_flow_field(path)
    out = np.zeros(N_TABLES * n_steps * GRID, np.float64)
    lib().orc_build_tables_f64(u, v, b, int(n_steps), 4.0 * float(Ne) * float(rho),
                               4.0 * float(Ne) * float(mu), out)
    return out.reshape(N_TABLES, n_steps, GRID)


def ref_tables(path, n_steps, Ne, mu, rho):
    """Tables from the reference's own ``build_flow_field_cache`` (oracle/_ref)."""
    srho, smu = scaled_rates(Ne, mu, rho)
    out = np.zeros(N_TABLES * n_steps * GRID, np.float32)
    b = np.zeros(4, np.float32)
    rc = ref_lib().ref_build_tables(os.fsencode(path), int(n_steps), float(srho), float(smu), out, b)
    if rc != 0:
        raise ValueError("reference table build failed")
    return out.reshape(N_TABLES, n_steps, GRID)


# ---------------------------------------------------------------- walk

def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def fb_window_f32(G, positions, pairs, tables, Ne, site_start=0, block_S=None,
                  ci=False, posterior=False):
    """FP32 forward-backward over ``[site_start, site_start + block_S)``.

    Returns a dict of site-major ``(block_S, P)`` float32 arrays: ``mean``,
    plus ``lower``/``upper`` when ``ci`` and ``posterior_alpha``/``_beta``
    when ``posterior``.
    """
    G = np.ascontiguousarray(G, np.uint8)
    pos = np.ascontiguousarray(positions, np.float64)
    n, S = G.shape
    if block_S is None:
        block_S = S - site_start
    pi = np.ascontiguousarray([p[0] for p in pairs], np.int32)
    pj = np.ascontiguousarray([p[1] for p in pairs], np.int32)
    P = len(pairs)
    T = np.ascontiguousarray(tables, np.float32)
    N = T.shape[1]
    out = {"mean": np.zeros((block_S, P), np.float32)}
    if ci:
        out["lower"] = np.zeros((block_S, P), np.float32)
        out["upper"] = np.zeros((block_S, P), np.float32)
    if posterior:
        out["posterior_alpha"] = np.zeros((block_S, P), np.float32)
        out["posterior_beta"] = np.zeros((block_S, P), np.float32)
    rc = lib().orc_fb_window_f32(G, n, S, pos, pi if P else np.zeros(1, np.int32),
                                 pj if P else np.zeros(1, np.int32), P, T.ravel(), N,
                                 float(np.float32(Ne)), int(site_start), int(block_S), out["mean"].reshape(-1) if out["mean"].size else np.zeros(1, np.float32),
                                 _ptr(out.get("lower")), _ptr(out.get("upper")),
                                 _ptr(out.get("posterior_alpha")), _ptr(out.get("posterior_beta")))
    if rc != 0:
        raise ValueError(f"oracle fb_window_f32 failed ({rc})")
    return out


def fb_window_f64(G, positions, pairs, tables64, Ne, site_start=0, block_S=None, exact_pow10=False):
    """FP64 restatement; returns ``mean``, ``posterior_alpha``, ``posterior_beta`` (float64)."""
    G = np.ascontiguousarray(G, np.uint8)
    pos = np.ascontiguousarray(positions, np.float64)
    n, S = G.shape
    if block_S is None:
        block_S = S - site_start
    pi = np.ascontiguousarray([p[0] for p in pairs], np.int32)
    pj = np.ascontiguousarray([p[1] for p in pairs], np.int32)
    P = len(pairs)
    T = np.ascontiguousarray(tables64, np.float64)
    mean = np.zeros((block_S, P))
    a = np.zeros((block_S, P))
    b = np.zeros((block_S, P))
    rc = lib().orc_fb_window_f64(G, n, S, pos, pi, pj, P, T.ravel(), T.shape[1], float(Ne),
                                 int(bool(exact_pow10)), int(site_start), int(block_S),
                                 mean.reshape(-1), _ptr(a), _ptr(b))
    if rc != 0:
        raise ValueError(f"oracle fb_window_f64 failed ({rc})")
    return {"mean": mean, "posterior_alpha": a, "posterior_beta": b}


def h2_tables(tables):
    """The forward-only path's tables from the FP32 set: the composed
    transitions cache[k] (tables 0 and 1) rounded to FP16, as the reference
    stores them (``bindings.cpp:2098-2107``), widened back to float32."""
    T = np.asarray(tables, np.float32)
    return (np.ascontiguousarray(T[0].astype(np.float16).astype(np.float32)),
            np.ascontiguousarray(T[1].astype(np.float16).astype(np.float32)))


def fwd_only_f32(G, positions, pairs, h2, Ne, ci=False):
    """Forward-only filter (``run_fwd``): ``mean`` and, with ``ci``, ``lower`` /
    ``upper``, site-major float32 ``(S, P)``."""
    G = np.ascontiguousarray(G, np.uint8)
    pos = np.ascontiguousarray(positions, np.float64)
    n, S = G.shape
    pi = np.ascontiguousarray([p[0] for p in pairs], np.int32)
    pj = np.ascontiguousarray([p[1] for p in pairs], np.int32)
    P = len(pairs)
    tm, tc = h2
    out = {"mean": np.zeros((S, P), np.float32)}
    if ci:
        out["lower"] = np.zeros((S, P), np.float32)
        out["upper"] = np.zeros((S, P), np.float32)
    if P and S:
        rc = lib().orc_fwd_only_f32(G, n, S, pos, pi, pj, P, tm.ravel(), tc.ravel(), tm.shape[0],
                                    float(np.float32(Ne)), out["mean"].reshape(-1), _ptr(out.get("lower")),
                                    _ptr(out.get("upper")))
        if rc != 0:
            raise ValueError(f"oracle fwd_only_f32 failed ({rc})")
    return out


def fwd_only_f64(G, positions, pairs, tables64, Ne):
    """FP64 forward-only filter on FP64 tables (no FP16 rounding)."""
    G = np.ascontiguousarray(G, np.uint8)
    pos = np.ascontiguousarray(positions, np.float64)
    n, S = G.shape
    pi = np.ascontiguousarray([p[0] for p in pairs], np.int32)
    pj = np.ascontiguousarray([p[1] for p in pairs], np.int32)
    T = np.asarray(tables64, np.float64)
    mean = np.zeros((S, len(pairs)))
    rc = lib().orc_fwd_only_f64(G, n, S, pos, pi, pj, len(pairs), np.ascontiguousarray(T[0]).ravel(),
                                np.ascontiguousarray(T[1]).ravel(), T.shape[1], float(Ne), mean.reshape(-1))
    if rc != 0:
        raise ValueError(f"oracle fwd_only_f64 failed ({rc})")
    return {"mean": mean}


def log_codes(mean, bits, Ne):
    """Log-quantized codes of FP32 means ``(S, P)`` (``gamma_smc.cu:23-39``):
    uint8 ``(S, P)``, or for 4 bits ``(ceil(S / 2), P)`` with the even site
    in the low nibble (``gamma_smc.cu:185-205``)."""
    mean = np.ascontiguousarray(mean, np.float32)
    inv = np.float32(1.0) / np.float32(np.log(np.float32(20.0) * np.float32(Ne)))
    levels = 255.0 if bits == 8 else 15.0
    q = np.zeros(mean.shape, np.uint8)
    if mean.size:
        lib().orc_log_codes(mean.reshape(-1), mean.size, float(inv), levels, q.reshape(-1))
    if bits == 8:
        return q
    S, P = q.shape
    if S % 2:
        q = np.concatenate([q, np.zeros((1, P), np.uint8)], axis=0)
    return (q[0::2] | (q[1::2] << 4)).astype(np.uint8)


def forward_states_f32(G, positions, pair, tables):
    G = np.ascontiguousarray(G, np.uint8)
    pos = np.ascontiguousarray(positions, np.float64)
    n, S = G.shape
    T = np.ascontiguousarray(tables, np.float32)
    m = np.zeros(S, np.float32)
    c = np.zeros(S, np.float32)
    rc = lib().orc_forward_states_f32(G, n, S, pos, int(pair[0]), int(pair[1]), T.ravel(), T.shape[1], m, c)
    if rc != 0:
        raise ValueError("oracle forward_states failed")
    return m, c


def fast_pow10(x):
    return lib().orc_fast_pow10(float(x))


def seg_steps(delta):
    return lib().orc_seg_steps(float(delta))


def dev_bilinear(tm, tc, m, c):
    om, oc = ctypes.c_float(), ctypes.c_float()
    lib().orc_dev_bilinear(np.ascontiguousarray(tm, np.float32), np.ascontiguousarray(tc, np.float32),
                           float(m), float(c), ctypes.byref(om), ctypes.byref(oc))
    return om.value, oc.value


def entropy_table():
    f = np.zeros(5001, np.float32)
    d = np.zeros(5001, np.float64)
    lib().orc_entropy_table(f, d)
    return f, d


# ---------------------------------------------------------------- host logic
# Restatements of python/gamma_smc_cu/infer.py and src/bindings.cpp.

def filter_segregating(G, positions):
    """``infer.py:102-126``: keep sites with 0 < allele count < n."""
    n = G.shape[0]
    ac = G.sum(axis=0, dtype=np.int64)
    keep = (ac > 0) & (ac < n)
    return np.ascontiguousarray(G[:, keep]), np.ascontiguousarray(positions[keep]), keep


def estimate_scaled_params(G, positions, mu, rho, Ne):
    """``infer.py:44-99``: per-diploid heterozygosity over [0, last_pos + 1)."""
    n, S = G.shape
    if n < 2 or S == 0 or n % 2:
        return float(mu), float(rho)
    L = float(positions[-1] + 1.0)
    if L <= 0:
        return float(mu), float(rho)
    het = (G[0::2] != G[1::2]).sum(axis=1, dtype=np.int64)
    pi_hat = float(het.mean()) / L
    if not (pi_hat > 0 and math.isfinite(pi_hat)):
        return float(mu), float(rho)
    ratio = float(rho) / max(float(mu), 1e-30)
    return pi_hat / (4.0 * float(Ne)), pi_hat * ratio / (4.0 * float(Ne))


def all_pairs(n):
    """``infer.py:272-273``: (i, j) for i in range(n) for j < i."""
    return [(i, j) for i in range(n) for j in range(i)]


def pair_to_ij(p):
    """``include/gamma_smc_cu/types.h:56-62``."""
    r = (1.0 + math.sqrt(1.0 + 8.0 * p)) / 2.0
    i = int(r)
    if i * (i - 1) // 2 > p:
        i -= 1
    return i, p - i * (i - 1) // 2


def ij_to_pair(i, j):
    """``include/gamma_smc_cu/types.h:64-68``."""
    if i < j:
        i, j = j, i
    return i * (i - 1) // 2 + j


def block_windows(S, core, flank):
    """``bindings.cpp:2517-2534``: (core_start, core_stop, padded_start, padded_stop)."""
    out = []
    for cs in range(0, S, core):
        ce = min(S, cs + core)
        out.append((cs, ce, max(0, cs - flank), min(S, ce + flank)))
    return np.array(out, dtype=np.int32).reshape(-1, 4)


def recommend_core_block_size(n_sites, n_pairs, mean_only, flank_sites, free_bytes):
    """``infer.py:151-214``."""
    if n_pairs <= 0:
        return min(n_sites, 32768), 0
    bps = 4 * (2 + (1 if mean_only else 3))
    headroom = max(512 * 1024 * 1024, int(free_bytes * 0.20))
    budget = max(0, free_bytes - headroom)
    core_max = budget // (bps * n_pairs) - 2 * flank_sites
    if core_max <= 0:
        chosen = min(n_sites, 1024)
    elif core_max >= n_sites:
        chosen = n_sites
    else:
        chosen = max(1024, min(1 << int(math.log2(core_max)), 32768, n_sites))
    padded = chosen + 2 * flank_sites
    return chosen, bps * padded * n_pairs


def fb_blockwise_f32(G, positions, pairs, tables, Ne, core, flank, ci=False, posterior=False):
    """Blockwise decode: each padded window decoded on its own, core rows
    stitched into the site-major output (``bindings.cpp:2991-3190``)."""
    S = G.shape[1]
    P = len(pairs)
    blocks = block_windows(S, core, flank)
    keys = ["mean"] + (["lower", "upper"] if ci else []) + (["posterior_alpha", "posterior_beta"] if posterior else [])
    out = {k: np.zeros((S, P), np.float32) for k in keys}
    for cs, ce, ps, pe in blocks:
        r = fb_window_f32(G, positions, pairs, tables, Ne, ps, pe - ps, ci, posterior)
        for k in keys:
            out[k][cs:ce] = r[k][cs - ps:ce - ps]
    out["blocks"] = blocks
    return out


def fb_blockwise_f64(G, positions, pairs, tables64, Ne, core, flank, exact_pow10=False):
    """FP64 blockwise decode (same windows as :func:`fb_blockwise_f32`);
    returns the float64 ``mean``."""
    S = G.shape[1]
    out = np.zeros((S, len(pairs)))
    for cs, ce, ps, pe in block_windows(S, core, flank):
        r = fb_window_f64(G, positions, pairs, tables64, Ne, ps, pe - ps, exact_pow10)
        out[cs:ce] = r["mean"][cs - ps:ce - ps]
    return {"mean": out}


def site_summary(mean):
    """Per-site reduction of an FP32 ``(S, P)`` posterior mean over pairs
    (the summary path's arithmetic; ``gamma_smc_flow.cu:1574-1683`` sums with
    float atomics in no fixed order, this fixes one). Pairs go in groups of
    64 (absent pairs 0); in each group, four parts of 16 consecutive pairs
    are summed in order in FP64 from 0.0 and joined as
    ``(part0 + part1) + (part2 + part3)``; the groups are added in order
    from 0.0; mean = float(sum / P). Min and max are exact."""
    mean = np.asarray(mean, np.float32)
    S, P = mean.shape
    G = -(-P // 64)
    x = np.zeros((S, G * 64), np.float64)
    x[:, :P] = mean
    x = x.reshape(S, G, 4, 16)
    t = np.zeros((S, G, 4), np.float64)
    for i in range(16):
        t = t + x[..., i]
    grp = (t[..., 0] + t[..., 1]) + (t[..., 2] + t[..., 3])
    tot = np.zeros(S, np.float64)
    for g in range(G):
        tot = tot + grp[:, g]
    return {"site_mean": (tot / P).astype(np.float32), "site_min": mean.min(axis=1), "site_max": mean.max(axis=1)}


GENE_BIN_EDGES = np.linspace(np.log(10.0), np.log(1_000_000.0), 51)  # infer_chromosome.py:52-58


def gene_site_indices(positions, genes):
    """Site indices of each ``(start, end)`` gene: ``start <= position <=
    end`` (``analysis/genome_wide/infer_chromosome.py:101-112``)."""
    positions = np.asarray(positions)
    return [np.where((positions >= s) & (positions <= e))[0].astype(np.int64) for s, e in genes]


def gene_pair_values(mean, positions, genes):
    """Per gene (``None`` below 2 sites), the per-pair float32 values of
    ``infer_chromosome.py:184-207``: the mean over the gene's sites of
    ``max(mean, 1)`` and of ``log(max(mean, 1))``, both float32 as numpy
    computes them there."""
    mean_safe = np.maximum(np.asarray(mean, np.float32), 1.0)
    log_mean = np.log(mean_safe)
    out = []
    for idxs in gene_site_indices(positions, genes):
        if idxs.size < 2:
            out.append(None)
            continue
        out.append((mean_safe[idxs, :].mean(axis=0), log_mean[idxs, :].mean(axis=0)))
    return out


def gene_accumulators(mean, positions, genes, edges=GENE_BIN_EDGES, acc=None):
    """One chunk's update of the per-gene accumulators
    (``infer_chromosome.py:157-227``): ``mean`` is the chunk's ``(S, P)``
    float32 posterior mean, ``genes`` ``(start, end)`` bp pairs. Adds into
    ``acc`` (fresh accumulators when ``None``) and returns it."""
    edges = np.asarray(edges, np.float64)
    nb = edges.size - 1
    n_genes = len(genes)
    if acc is None:
        acc = {"count": np.zeros(n_genes, np.int64), "lin_sum": np.zeros(n_genes, np.float64),
               "log_sum": np.zeros(n_genes, np.float64), "log_sq_sum": np.zeros(n_genes, np.float64),
               "min_lin": np.full(n_genes, np.inf, np.float64), "min_log": np.full(n_genes, np.inf, np.float64),
               "histogram": np.zeros((n_genes, nb), np.int64)}
    n_pairs = np.asarray(mean).shape[1]
    for gi, v in enumerate(gene_pair_values(mean, positions, genes)):
        if v is None:
            continue
        lin, lg = v
        acc["count"][gi] += n_pairs
        acc["lin_sum"][gi] += lin.sum()
        acc["log_sum"][gi] += lg.sum()
        acc["log_sq_sum"][gi] += (lg * lg).sum()
        if lin.min() < acc["min_lin"][gi]:
            acc["min_lin"][gi] = lin.min()
        if lg.min() < acc["min_log"][gi]:
            acc["min_log"][gi] = lg.min()
        bins = np.clip(np.digitize(lg, edges) - 1, 0, nb - 1)
        np.add.at(acc["histogram"][gi], bins, 1)
    return acc


def log_error_report(mean32, mean64):
    """|log mean32 - log mean64| summary (max, p99, median, mean)."""
    e = np.abs(np.log(np.asarray(mean32, np.float64)) - np.log(np.asarray(mean64, np.float64)))
    return {"max": float(e.max()), "p99": float(np.percentile(e, 99)), "median": float(np.median(e)),
            "mean": float(e.mean()), "n": int(e.size)}


def bitpack(G):
    """Bit b of word w of row h is ``G[h, 64 w + b] & 1``; tail bits zero
    (``bitpack.cu:11-30``)."""
    G = np.asarray(G, np.uint8)
    n, S = G.shape
    W = (S + 63) // 64
    pad = np.zeros((n, W * 64), np.uint8)
    pad[:, :S] = G & 1
    bits = pad.reshape(n, W, 64).astype(np.uint64)
    return (bits << np.arange(64, dtype=np.uint64)).sum(axis=2, dtype=np.uint64)


def unpack(packed, n, S):
    packed = np.asarray(packed, np.uint64)
    bits = (packed[:, :, None] >> np.arange(64, dtype=np.uint64)) & np.uint64(1)
    return bits.reshape(n, -1)[:, :S].astype(np.uint8)
