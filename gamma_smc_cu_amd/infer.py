"""``infer`` and ``infer_blockwise``: the drop-in Python API.

Same signatures, defaults, validation messages, warnings and result layout
as the reference's ``python/gamma_smc_cu/infer.py:217-484``. The native work
runs through ``gamma_smc_cu_amd._core`` (HIP kernels on the current device),
looked up at call time so it can be replaced in tests.

Result dict: ``mean`` float32 ``(n_sites, n_pairs)`` site-major (also under
``tmrca``), ``lower``/``upper`` when ``mean_only`` is false,
``posterior_alpha``/``posterior_beta`` when ``return_posterior``, ``pairs``
(list of int tuples), ``positions`` (float64, segregating sites only), and
``blocks`` int32 ``(n_blocks, 4)`` for the blockwise form.
"""

from __future__ import annotations

import math
import os
import warnings

import numpy as np

_BYTES_PER_FLOAT32 = 4
_MIN_BLOCK_SITES = 1024
_MAX_BLOCK_SITES = 32768
_DEFAULT_BLOCK_SITES = 8192


def _coerce_inputs(G_or_ts, positions):
    """Tree sequence or matrix to contiguous ``(G uint8, positions float64)``
    (``infer.py:20-41``)."""
    if hasattr(G_or_ts, "genotype_matrix"):
        ts = G_or_ts
        G = ts.genotype_matrix().T.astype(np.uint8, copy=False)
        pos = np.array([var.position for var in ts.variants()], dtype=np.float64)
        return np.ascontiguousarray(G), pos
    if positions is None:
        raise ValueError("positions is required when G_or_ts is a genotype matrix.")
    G = np.ascontiguousarray(G_or_ts, dtype=np.uint8)
    if G.ndim != 2:
        raise ValueError("G_or_ts must be a 2D haplotype matrix or a tree sequence.")
    pos = np.ascontiguousarray(positions, dtype=np.float64)
    if pos.ndim != 1:
        raise ValueError("positions must be a 1D array of site coordinates.")
    if G.shape[1] != pos.shape[0]:
        raise ValueError("positions length must match the number of sites in G_or_ts.")
    return G, pos


def _estimate_scaled_params(G, positions, mu, rho, Ne):
    """Data-implied ``(mu, rho)`` as gamma_smc's auto-estimation mode forms
    them (``infer.py:44-99``).

    pi_hat = mean over diploid individuals (rows 2k, 2k+1) of the number of
    sites where their two haplotypes differ, over the span
    ``[0, positions[-1] + 1)``; then ``4 Ne mu_eff = pi_hat`` and
    ``4 Ne rho_eff = pi_hat * rho / mu``. FP64 throughout. Inputs with fewer
    than two or an odd number of haplotypes, no sites, a non-positive span or
    a non-positive pi_hat keep ``(mu, rho)``.
    """
    n, S = G.shape
    if n < 2 or S == 0 or n % 2 != 0:
        return float(mu), float(rho)
    span = float(positions[-1] + 1.0)
    if span <= 0:
        return float(mu), float(rho)
    het = np.count_nonzero(G[0::2] != G[1::2], axis=1).astype(np.int64)
    pi_hat = float(het.mean()) / span
    if not (pi_hat > 0 and math.isfinite(pi_hat)):
        return float(mu), float(rho)
    ratio = float(rho) / max(float(mu), 1e-30)
    return pi_hat / (4.0 * float(Ne)), pi_hat * ratio / (4.0 * float(Ne))


def _estimate_scaled_params_packed(packed, n_haps, positions, mu, rho, Ne):
    """:func:`_estimate_scaled_params` on bitpacked rows (``packed[h][w]``
    bit ``s & 63`` of word ``s >> 6``, zero tail bits): the per-diploid
    heterozygous site counts are popcounts of ``row(2k) ^ row(2k+1)``."""
    S = int(positions.shape[0])
    if n_haps < 2 or S == 0 or n_haps % 2 != 0:
        return float(mu), float(rho)
    span = float(positions[-1] + 1.0)
    if span <= 0:
        return float(mu), float(rho)
    rows = np.asarray(packed, dtype=np.uint64)
    het = np.bitwise_count(rows[0:n_haps:2] ^ rows[1:n_haps:2]).sum(axis=1, dtype=np.int64)
    pi_hat = float(het.mean()) / span
    if not (pi_hat > 0 and math.isfinite(pi_hat)):
        return float(mu), float(rho)
    ratio = float(rho) / max(float(mu), 1e-30)
    return pi_hat / (4.0 * float(Ne)), pi_hat * ratio / (4.0 * float(Ne))


def _filter_segregating(G, positions):
    """Drop sites whose allele count is 0 or n (``infer.py:102-126``); no copy
    when every site segregates."""
    n = G.shape[0]
    ac = G.sum(axis=0, dtype=np.int64)
    keep = (ac > 0) & (ac < n)
    if bool(keep.all()):
        return G, positions, keep
    return np.ascontiguousarray(G[:, keep]), np.ascontiguousarray(positions[keep]), keep


def _resolve_flow_field_path(flow_field_path):
    """The caller's path, else the packaged field (``infer.py:129-144``)."""
    if flow_field_path is not None:
        return flow_field_path
    from .flow_field import DEFAULT_PATH

    if os.path.exists(DEFAULT_PATH):
        return DEFAULT_PATH
    raise FileNotFoundError("No flow field file found. Pass flow_field_path explicitly.")


def _normalize_pairs(pairs):
    return [(int(i), int(j)) for i, j in pairs]


def _gpu_bytes_per_pair_site(mean_only):
    """Per pair and padded site, as the reference sizes blocks
    (``infer.py:151-154``): two forward floats plus the output arrays."""
    return _BYTES_PER_FLOAT32 * (2 + (1 if mean_only else 3))


def _query_free_gpu_bytes():
    try:
        from . import _core

        free_bytes, _total = _core.cuda_mem_info()
        return int(free_bytes)
    except Exception:
        return None


def _estimate_block_gpu_bytes(core_block_sites, flank_sites, n_pairs, mean_only, n_sites=None):
    padded = core_block_sites + 2 * flank_sites
    if n_sites is not None:
        padded = min(padded, n_sites)
    return _gpu_bytes_per_pair_site(mean_only) * padded * n_pairs


def _recommend_core_block_size(n_sites, n_pairs, mean_only, flank_sites, free_bytes, headroom_fraction=0.20,
                               min_headroom_bytes=512 * 1024 * 1024, min_block=_MIN_BLOCK_SITES,
                               max_block=_MAX_BLOCK_SITES):
    """Largest power-of-two core that fits all pairs in one batch, clamped to
    ``[min_block, max_block]``, or every site when they all fit
    (``infer.py:181-214``)."""
    if n_pairs <= 0:
        return min(n_sites, max_block), 0
    bps = _gpu_bytes_per_pair_site(mean_only)
    headroom = max(min_headroom_bytes, int(free_bytes * headroom_fraction))
    budget = max(0, free_bytes - headroom)
    core_max = budget // (bps * n_pairs) - 2 * flank_sites
    if core_max <= 0:
        chosen = min(n_sites, min_block)
    elif core_max >= n_sites:
        chosen = n_sites
    else:
        chosen = max(min_block, min(1 << int(math.log2(core_max)), max_block, n_sites))
    return chosen, _estimate_block_gpu_bytes(chosen, flank_sites, n_pairs, mean_only)


def _finish(result, pairs, positions):
    result["pairs"] = pairs
    result["positions"] = positions
    result["tmrca"] = result["mean"]
    return result


def infer(G_or_ts, positions=None, mu=1.25e-8, rho=1e-8, Ne=10000, pairs=None, flow_field_path=None,
          mean_only=True, return_posterior=False, auto_estimate_theta=True):
    """Posterior TMRCA (generations) at every segregating site for each pair.

    Drop-in for ``gamma_smc_cu.infer`` (``infer.py:217-296``): monomorphic
    sites are dropped first; ``pairs`` defaults to all ``(i, j)`` with
    ``j < i`` in the order ``(1,0), (2,0), (2,1), ...``; with
    ``auto_estimate_theta`` the scaled rates come from the data. Returns the
    dict described in the module docstring.
    """
    from . import _core

    G, positions = _coerce_inputs(G_or_ts, positions)
    G, positions, _ = _filter_segregating(G, positions)
    n = G.shape[0]
    if pairs is None:
        pairs = [(i, j) for i in range(n) for j in range(i)]
    else:
        pairs = _normalize_pairs(pairs)
    flow_field_path = _resolve_flow_field_path(flow_field_path)
    if auto_estimate_theta:
        kmu, krho = _estimate_scaled_params(G, positions, mu, rho, Ne)
    else:
        kmu, krho = float(mu), float(rho)
    ctx = _core.FlowContext(G, positions, float(Ne), kmu, krho, flow_field_path, 0)
    result = ctx.run_fb(pairs, mean_only=mean_only, return_posterior=return_posterior)
    return _finish(result, pairs, positions)


def infer_blockwise(G_or_ts, positions=None, mu=1.25e-8, rho=1e-8, Ne=10000, pairs=None, flow_field_path=None,
                    mean_only=True, core_block_sites="auto", flank_sites=2048, pair_batch_size=-1, max_streams=1,
                    verbose=False, return_posterior=False, auto_estimate_theta=True):
    """Blockwise decoding: padded site windows decoded from the prior, core
    rows stitched into ``(n_sites, n_pairs)``.

    Drop-in for ``gamma_smc_cu.infer_blockwise`` (``infer.py:299-484``),
    including its argument checks, ``core_block_sites="auto"`` sizing from
    free device memory, and the ``flank_sites=0`` multi-block rejection.
    """
    from . import _core

    if pairs is None:
        raise ValueError("infer_blockwise requires explicit pairs in v1.")
    if return_posterior and max_streams > 1:
        raise ValueError(
            "infer_blockwise: return_posterior=True is currently only supported "
            "with max_streams=1. Run blockwise posterior decoding in single-stream mode.")
    if flank_sites < 0:
        raise ValueError("flank_sites must be non-negative.")
    if pair_batch_size == 0 or pair_batch_size < -1:
        raise ValueError("pair_batch_size must be positive or -1 for auto.")
    if max_streams <= 0:
        raise ValueError("max_streams must be positive.")
    auto_block = core_block_sites in ("auto", None)
    if not auto_block:
        if not isinstance(core_block_sites, (int, np.integer)):
            raise TypeError(
                f"core_block_sites must be int or 'auto', got {type(core_block_sites).__name__}.")
        if core_block_sites <= 0:
            raise ValueError("core_block_sites must be positive (or 'auto').")

    G, positions = _coerce_inputs(G_or_ts, positions)
    G, positions, _ = _filter_segregating(G, positions)
    pairs = _normalize_pairs(pairs)
    n_sites = G.shape[1]
    n_pairs = len(pairs)

    free_bytes = _query_free_gpu_bytes()
    if auto_block:
        if free_bytes is None:
            warnings.warn(
                "infer_blockwise(core_block_sites='auto') could not query GPU "
                f"memory; falling back to core_block_sites={_DEFAULT_BLOCK_SITES}.",
                stacklevel=2)
            core_block_sites = _DEFAULT_BLOCK_SITES
        else:
            core_block_sites, est = _recommend_core_block_size(n_sites, n_pairs, mean_only, flank_sites, free_bytes)
            if verbose:
                print(f"[infer_blockwise] auto core_block_sites={core_block_sites} "
                      f"(free_gpu={free_bytes / 1e9:.2f}GB, est_block={est / 1e9:.2f}GB, "
                      f"n_blocks={max(1, math.ceil(n_sites / max(core_block_sites, 1)))})", flush=True)
    elif free_bytes is not None and n_pairs > 0:
        est = _estimate_block_gpu_bytes(core_block_sites, flank_sites, n_pairs, mean_only, n_sites=n_sites)
        if est > free_bytes:
            rec, _ = _recommend_core_block_size(n_sites, n_pairs, mean_only, flank_sites, free_bytes)
            warnings.warn(
                f"infer_blockwise: core_block_sites={core_block_sites} with flank_sites={flank_sites} "
                f"for {n_pairs} pairs needs ~{est / 1e9:.2f}GB GPU memory but only {free_bytes / 1e9:.2f}GB "
                f"is free. The native backend will fall back to chunking pairs (slower). Suggested: "
                f"core_block_sites={rec} (or pass 'auto').",
                stacklevel=2)

    if flank_sites == 0 and core_block_sites < n_sites:
        raise ValueError(
            "infer_blockwise: flank_sites=0 with core_block_sites < n_sites "
            f"(core_block_sites={core_block_sites}, n_sites={n_sites}) produces invalid results because "
            "each block restarts the forward/backward pass from the marginal prior with no warm-up history. "
            "Pass flank_sites>=256 (default 2048) or set core_block_sites>=n_sites to decode the whole "
            "sequence as a single block.")

    batch = pair_batch_size
    if batch == -1:
        batch = max(1, n_pairs)

    flow_field_path = _resolve_flow_field_path(flow_field_path)
    if auto_estimate_theta:
        kmu, krho = _estimate_scaled_params(G, positions, mu, rho, Ne)
    else:
        kmu, krho = float(mu), float(rho)
    ctx = _core.FlowContext(G, positions, float(Ne), kmu, krho, flow_field_path, 0)
    result = ctx.run_fb_blockwise(pairs, core_block_sites=int(core_block_sites), flank_sites=flank_sites,
                                  pair_batch_size=batch, max_streams=max_streams, mean_only=mean_only,
                                  return_posterior=return_posterior)
    return _finish(result, pairs, positions)
