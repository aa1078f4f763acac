"""Native surface of the decoder: ctypes bindings to ``libgsmc.so``.

Mirrors the attributes the reference's pybind11 module ``gamma_smc_cu._core``
exposes on this path (``src/bindings.cpp:3497-3731``): ``FlowContext`` with
``run_fb`` / ``run_fb_blockwise`` and the ``n_haps`` / ``n_sites`` /
``device_id`` properties, ``cuda_mem_info``, ``bitpack`` and ``unpack``.
``infer`` / ``infer_blockwise`` look these up as attributes at call time, so
tests can replace them (the reference's own tests monkeypatch
``_core.FlowContext`` and ``_core.cuda_mem_info`` the same way).

The library is loaded on first use. There is no CPU fallback: without the
built library, or without a GPU, every call raises.
"""

from __future__ import annotations

import ctypes
import os
import threading

import numpy as np

from .flow_field import DEFAULT_PATH

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libgsmc.so")

GSMC_CI = 1
GSMC_POSTERIOR = 2
GSMC_SUMMARY = 4
GSMC_Q8 = 8
GSMC_Q4 = 16
_E_INVAL, _E_HIP, _E_NOMEM, _E_IO = -1, -2, -3, -4

# Every symbol include/gamma_smc_amd.h declares.
EXPORTS = (
    "gsmc_last_error", "gsmc_build_tables", "gsmc_device_count", "gsmc_set_device", "gsmc_mem_info", "gsmc_bitpack",
    "gsmc_unpack", "gsmc_context_create", "gsmc_context_create_packed", "gsmc_context_destroy",
    "gsmc_context_info", "gsmc_context_set_tuning", "gsmc_run_fb", "gsmc_block_windows",
    "gsmc_run_fb_blockwise", "gsmc_run_fb_device", "gsmc_malloc", "gsmc_free", "gsmc_memcpy_htod",
    "gsmc_memcpy_dtoh", "gsmc_memset", "gsmc_synchronize", "gsmc_stream_create", "gsmc_stream_destroy",
    "gsmc_stream_synchronize", "gsmc_event_create", "gsmc_event_destroy", "gsmc_event_record",
    "gsmc_event_elapsed_ms", "gsmc_batch_create", "gsmc_batch_destroy", "gsmc_batch_launch", "gsmc_batch_wait",
    "gsmc_batch_query", "gsmc_checkpoint_bytes", "gsmc_device_copy", "gsmc_memcpy_htod_async",
    "gsmc_memcpy_dtoh_async", "gsmc_host_alloc", "gsmc_host_free", "gsmc_event_query",
    "gsmc_batch_reserve", "gsmc_run_fb_summary", "gsmc_run_fwd", "gsmc_run_fwd_device", "gsmc_batch_capture",
    "gsmc_graph_launch", "gsmc_graph_destroy", "gsmc_context_set_state_budget", "gsmc_context_release_buffers",
    "gsmc_run_fwd_states", "gsmc_get_device", "gsmc_batch_buffer_bytes", "gsmc_run_fb_device_noalloc",
    "gsmc_lookup_probe", "gsmc_run_fb_gene_stats",
)


class Job(ctypes.Structure):
    """``gsmc_job``: one job of a batched launch (device pointers)."""

    _fields_ = [("site_lo", ctypes.c_int64), ("site_hi", ctypes.c_int64), ("core_block_sites", ctypes.c_int64),
                ("flank_sites", ctypes.c_int64), ("pair_i", ctypes.c_void_p), ("pair_j", ctypes.c_void_p),
                ("n_pairs", ctypes.c_int64), ("mean", ctypes.c_void_p), ("lower", ctypes.c_void_p),
                ("upper", ctypes.c_void_p), ("alpha", ctypes.c_void_p), ("beta", ctypes.c_void_p),
                ("site_mean", ctypes.c_void_p), ("site_min", ctypes.c_void_p), ("site_max", ctypes.c_void_p)]

_lib = None
_lock = threading.Lock()

_vp = ctypes.c_void_p
_i64 = ctypes.c_int64
_i32 = ctypes.c_int


def _declare(L):
    L.gsmc_last_error.restype = ctypes.c_char_p
    L.gsmc_last_error.argtypes = []
    sig = {
        "gsmc_build_tables": [ctypes.c_char_p, _i32, ctypes.c_double, ctypes.c_double, ctypes.c_double, _i32,
                              _vp, _vp, _vp],
        "gsmc_device_count": [ctypes.POINTER(ctypes.c_int)],
        "gsmc_set_device": [_i32],
        "gsmc_get_device": [ctypes.POINTER(ctypes.c_int)],
        "gsmc_lookup_probe": [_vp, _i64, _i32, _vp, _vp, ctypes.POINTER(ctypes.c_float)],
        "gsmc_run_fwd_states": [_vp, _vp, _vp, _i64, _vp, _vp],
        "gsmc_mem_info": [ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_uint64)],
        "gsmc_bitpack": [_vp, _i64, _i64, _vp],
        "gsmc_unpack": [_vp, _i64, _i64, _vp],
        "gsmc_context_create": [_vp, _i64, _i64, _vp, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                ctypes.c_char_p, _i32, _i32, ctypes.POINTER(_vp)],
        "gsmc_context_create_packed": [_vp, _i64, _i64, _vp, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                       ctypes.c_char_p, _i32, _i32, ctypes.POINTER(_vp)],
        "gsmc_context_info": [_vp, ctypes.POINTER(_i64), ctypes.POINTER(_i64), ctypes.POINTER(ctypes.c_int),
                              ctypes.POINTER(ctypes.c_float)],
        "gsmc_context_set_tuning": [_vp, _i32, _i32, _i32, _i32, _i64],
        "gsmc_context_set_state_budget": [_vp, ctypes.c_uint64],
        "gsmc_context_release_buffers": [_vp],
        "gsmc_run_fb": [_vp, _vp, _vp, _i64, _i32, _vp, _vp, _vp, _vp, _vp],
        "gsmc_block_windows": [_i64, _i64, _i64, _vp, _i64, ctypes.POINTER(_i64)],
        "gsmc_run_fb_blockwise": [_vp, _vp, _vp, _i64, _i64, _i64, _i64, _i32, _i32, _vp, _vp, _vp, _vp, _vp],
        "gsmc_run_fb_device": [_vp, _vp, _vp, _i64, _i64, _i64, _i32, _vp, _vp, _vp, _vp, _vp, _vp],
        "gsmc_run_fb_device_noalloc": [_vp, _vp, _vp, _i64, _i64, _i64, _i32, _vp, _vp, _vp, _vp, _vp, _vp],
        "gsmc_malloc": [ctypes.POINTER(_vp), ctypes.c_uint64],
        "gsmc_free": [_vp],
        "gsmc_memcpy_htod": [_vp, _vp, ctypes.c_uint64],
        "gsmc_memcpy_dtoh": [_vp, _vp, ctypes.c_uint64],
        "gsmc_memset": [_vp, _i32, ctypes.c_uint64],
        "gsmc_synchronize": [],
        "gsmc_stream_create": [ctypes.POINTER(_vp)],
        "gsmc_stream_destroy": [_vp],
        "gsmc_stream_synchronize": [_vp],
        "gsmc_event_create": [ctypes.POINTER(_vp)],
        "gsmc_event_destroy": [_vp],
        "gsmc_event_record": [_vp, _vp],
        "gsmc_event_elapsed_ms": [_vp, _vp, ctypes.POINTER(ctypes.c_float)],
        "gsmc_batch_create": [_vp, ctypes.POINTER(_vp)],
        "gsmc_batch_launch": [_vp, _vp, _i64, _i32, _vp],
        "gsmc_batch_wait": [_vp],
        "gsmc_batch_query": [_vp, ctypes.POINTER(ctypes.c_int)],
        "gsmc_checkpoint_bytes": [_vp, _i64, _i64, _i64, _i64, ctypes.POINTER(ctypes.c_uint64)],
        "gsmc_device_copy": [_vp, _vp, ctypes.c_uint64, _vp],
        "gsmc_memcpy_htod_async": [_vp, _vp, ctypes.c_uint64, _vp],
        "gsmc_memcpy_dtoh_async": [_vp, _vp, ctypes.c_uint64, _vp],
        "gsmc_host_alloc": [ctypes.POINTER(_vp), ctypes.c_uint64],
        "gsmc_host_free": [_vp],
        "gsmc_event_query": [_vp, ctypes.POINTER(ctypes.c_int)],
        "gsmc_batch_reserve": [_vp, ctypes.c_uint64, _i64, _i64],
        "gsmc_batch_buffer_bytes": [_vp, ctypes.POINTER(ctypes.c_uint64)],
        "gsmc_run_fb_summary": [_vp, _vp, _vp, _i64, _i64, _i64, _vp, _vp, _vp],
        "gsmc_run_fb_gene_stats": [_vp, _vp, _vp, _i64, _i64, _i64, _i64, _i64, _vp, _vp, _i32, _vp, _vp, _vp, _vp,
                                   _vp, _vp, _vp, _vp],
        "gsmc_run_fwd": [_vp, _vp, _vp, _i64, _i32, _vp, _vp, _vp],
        "gsmc_run_fwd_device": [_vp, _vp, _vp, _i64, _i32, _vp, _vp, _vp, _vp],
        "gsmc_batch_capture": [_vp, _vp, _i64, _i32, ctypes.POINTER(_vp)],
        "gsmc_graph_launch": [_vp, _vp],
    }
    for name, args in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = ctypes.c_int
    L.gsmc_context_destroy.argtypes = [_vp]
    L.gsmc_context_destroy.restype = None
    L.gsmc_batch_destroy.argtypes = [_vp]
    L.gsmc_batch_destroy.restype = None
    L.gsmc_graph_destroy.argtypes = [_vp]
    L.gsmc_graph_destroy.restype = None


def lib():
    """The loaded C library; raises if it has not been built."""
    global _lib
    if _lib is None:
        with _lock:
            if _lib is None:
                if not os.path.exists(LIB_PATH):
                    raise RuntimeError(
                        f"gamma_smc_cu_amd: native library {LIB_PATH} is not built; "
                        "run `python -c 'import __graft_entry__ as g; g.build()'` from the repository root")
                L = ctypes.CDLL(LIB_PATH)
                _declare(L)
                _lib = L
    return _lib


def check(rc):
    """Raise the Python exception the reference raises for the same failure."""
    if rc == 0:
        return
    msg = lib().gsmc_last_error().decode(errors="replace")
    if rc == _E_INVAL:
        raise ValueError(msg)
    if rc == _E_NOMEM:
        raise MemoryError(msg)
    raise RuntimeError(msg)


def _ptr(a):
    return None if a is None else ctypes.c_void_p(a.ctypes.data)


def cuda_mem_info():
    """``(free_bytes, total_bytes)`` of the current device (``bindings.cpp:3646-3652``)."""
    f, t = ctypes.c_uint64(), ctypes.c_uint64()
    check(lib().gsmc_mem_info(ctypes.byref(f), ctypes.byref(t)))
    return int(f.value), int(t.value)


def device_count():
    n = ctypes.c_int()
    check(lib().gsmc_device_count(ctypes.byref(n)))
    return n.value


def get_device_count():
    """Number of HIP devices (``bindings.cpp:3635-3639``)."""
    return device_count()


def set_device(device_id):
    """Make ``device_id`` the current device (``bindings.cpp:3630-3633``)."""
    check(lib().gsmc_set_device(int(device_id)))


def get_device():
    """The current device (``bindings.cpp:3641-3645``)."""
    d = ctypes.c_int()
    check(lib().gsmc_get_device(ctypes.byref(d)))
    return d.value


def bitpack(G):
    """Pack ``uint8 (n, S)`` rows into ``uint64 (n, ceil(S/64))`` on the device."""
    G = np.ascontiguousarray(G, dtype=np.uint8)
    if G.ndim != 2:
        raise ValueError("G must be 2-D (n_haplotypes, n_sites)")
    n, S = G.shape
    out = np.zeros((n, (S + 63) // 64), np.uint64)
    check(lib().gsmc_bitpack(_ptr(G), n, S, _ptr(out)))
    return out


def unpack(packed, n, S):
    """Inverse of :func:`bitpack`."""
    packed = np.ascontiguousarray(packed, dtype=np.uint64)
    out = np.zeros((int(n), int(S)), np.uint8)
    check(lib().gsmc_unpack(_ptr(packed), int(n), int(S), _ptr(out)))
    return out


def block_windows(n_sites, core_block_sites, flank_sites):
    """``(n_blocks, 4)`` int32 windows, as ``FlowContext::make_block_windows``."""
    nb = ctypes.c_int64()
    check(lib().gsmc_block_windows(int(n_sites), int(core_block_sites), int(flank_sites), None, 0,
                                   ctypes.byref(nb)))
    out = np.zeros((nb.value, 4), np.int32)
    check(lib().gsmc_block_windows(int(n_sites), int(core_block_sites), int(flank_sites), _ptr(out),
                                   nb.value, ctypes.byref(nb)))
    return out


def build_tables(flow_field_path, n_steps, Ne, mu, rho, table_precision="fp32"):
    """Host-built transition tables as the kernels read them:
    ``fwd``/``bwd`` float32 ``(n_steps, 2, 2550, 2)`` ([step][hom, het][node][m, c])
    and ``stretch`` ``(2550, 2)``."""
    fwd = np.zeros((n_steps, 2, 2550, 2), np.float32)
    bwd = np.zeros_like(fwd)
    stretch = np.zeros((2550, 2), np.float32)
    check(lib().gsmc_build_tables(os.fsencode(flow_field_path), int(n_steps), float(Ne), float(mu), float(rho),
                                  _PRECISION[table_precision], _ptr(fwd), _ptr(bwd), _ptr(stretch)))
    return fwd, bwd, stretch


def _split_pairs(pairs):
    if isinstance(pairs, np.ndarray) and pairs.ndim == 2 and pairs.shape[1] == 2 and pairs.dtype.kind in "iu":
        if pairs.size and (pairs.min() < np.iinfo(np.int32).min or pairs.max() > np.iinfo(np.int32).max):
            raise ValueError("pair index out of range")
        return (np.ascontiguousarray(pairs[:, 0], dtype=np.int32),
                np.ascontiguousarray(pairs[:, 1], dtype=np.int32))
    pairs = list(pairs)
    pi = np.ascontiguousarray([int(p[0]) for p in pairs], dtype=np.int32)
    pj = np.ascontiguousarray([int(p[1]) for p in pairs], dtype=np.int32)
    return pi, pj


_PRECISION = {"fp32": 0, "fp64": 1}


def check_pairs_host(pi, pj, n_haps):
    if pi.size and (pi.min() < 0 or pj.min() < 0 or pi.max() >= n_haps or pj.max() >= n_haps):
        raise ValueError("pair index out of range")


def dequantize(codes, bits, Ne):
    """Posterior means represented by log-quantized codes (the inverse of
    the code map, at the centre of each code's bin edge):
    ``exp(code / levels * log(20 Ne))``; uint4 codes are unpacked to one
    row per site first (``n_sites`` rows are returned for an even count;
    the caller drops the padding row of an odd count)."""
    codes = np.asarray(codes, np.uint8)
    if bits == 4:
        lo, hi = codes & 0x0F, codes >> 4
        codes = np.stack([lo, hi], axis=1).reshape(-1, codes.shape[1])
        levels = 15.0
    elif bits == 8:
        levels = 255.0
    else:
        raise ValueError("bits must be 4 or 8")
    log_max = float(np.log(np.float32(20.0) * np.float32(Ne)))
    return np.exp(codes.astype(np.float64) / levels * log_max).astype(np.float32)


class FlowContext:
    """Device context for flow-field Gamma-SMC decoding.

    Same constructor arguments and methods as the reference's
    ``_core.FlowContext`` (``bindings.cpp:2552-2670,3687-3730``), plus
    ``table_precision`` (``"fp32"``: reference-identical table build,
    ``"fp64"``: FP64 build) and ``packed=True`` to pass rows already
    bitpacked by :func:`bitpack`.
    """

    def __init__(self, G, positions, Ne=10000.0, mu=1.25e-8, rho=1e-8,
                 flow_field_path=DEFAULT_PATH, cache_steps=0, table_precision="fp32", packed=False,
                 n_sites=None):
        self._h = None
        pos = np.ascontiguousarray(positions, dtype=np.float64)
        if table_precision not in _PRECISION:
            raise ValueError("table_precision must be 'fp32' or 'fp64'")
        h = ctypes.c_void_p()
        path = os.fsencode(flow_field_path)
        if packed:
            rows = np.ascontiguousarray(G, dtype=np.uint64)
            n = rows.shape[0]
            S = int(n_sites) if n_sites is not None else pos.shape[0]
            if rows.shape[1] != (S + 63) // 64 or pos.shape[0] != S:
                raise ValueError("packed rows / positions do not match n_sites")
            check(lib().gsmc_context_create_packed(_ptr(rows), n, S, _ptr(pos), float(Ne), float(mu), float(rho),
                                                   path, int(cache_steps), _PRECISION[table_precision],
                                                   ctypes.byref(h)))
        else:
            G = np.ascontiguousarray(G, dtype=np.uint8)
            if G.ndim != 2:
                raise ValueError("G must be 2-D (n_haplotypes, n_sites)")
            n, S = G.shape
            if pos.shape[0] != S:
                raise ValueError("positions length must match the number of sites in G")
            check(lib().gsmc_context_create(_ptr(G), n, S, _ptr(pos), float(Ne), float(mu), float(rho), path,
                                            int(cache_steps), _PRECISION[table_precision], ctypes.byref(h)))
        self._h = h
        n_, S_, dev, ne = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int(), ctypes.c_float()
        check(lib().gsmc_context_info(h, ctypes.byref(n_), ctypes.byref(S_), ctypes.byref(dev), ctypes.byref(ne)))
        self._n, self._S, self._dev, self._Ne = n_.value, S_.value, dev.value, ne.value

    # properties (bindings.cpp:3727-3729)
    @property
    def n_haps(self):
        return self._n

    @property
    def n_sites(self):
        return self._S

    @property
    def device_id(self):
        return self._dev

    @property
    def handle(self):
        return self._h

    def close(self):
        if self._h is not None and _lib is not None:
            _lib.gsmc_context_destroy(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_tuning(self, checkpoint_k=8, pairs_per_lane=1, schedule="static", xcd_order=False,
                   min_waves_per_launch=8192):
        """Launch shape (results are identical under every setting).
        ``schedule``: ``"static"`` (workgroup g decodes item g),
        ``"persistent"`` (resident waves drain a device queue) or
        ``"dynamic"`` (one workgroup per item, each taking the next item of
        the queue).
        ``checkpoint_k=1`` stores the forward state at every core site
        (8 bytes per pair and site of device memory) instead of
        recomputing it in the backward kernel."""
        sched = {"static": 0, "persistent": 1, "dynamic": 2}[schedule]
        check(lib().gsmc_context_set_tuning(self._h, int(checkpoint_k), int(pairs_per_lane), sched,
                                            int(bool(xcd_order)), int(min_waves_per_launch)))

    def release_buffers(self):
        """Free the device workspaces this context keeps between calls (they
        keep their largest launch's buffers); the next call allocates again."""
        check(lib().gsmc_context_release_buffers(self._h))

    def set_state_budget(self, max_bytes=0):
        """Cap the device memory a launch uses for checkpoints / stored
        forward states (0: free memory less 2 GiB); larger launches run in
        consecutive phases. Results are identical under every setting."""
        check(lib().gsmc_context_set_state_budget(self._h, int(max_bytes)))

    def run_fb_codes(self, pairs, bits=8, core_block_sites=0, flank_sites=0):
        """Blockwise decode whose posterior mean comes back as log-quantized
        codes (``gamma_smc.cu:23-39``, ``bindings.cpp:1680-1810``): uint8
        ``(n_sites, n_pairs)`` for ``bits=8``; for ``bits=4`` two sites per
        byte, ``(ceil(n_sites / 2), n_pairs)``, even site in the low nibble.
        Codes map ``log(max(mean, 1))`` linearly from ``[log_min, log_max]``
        = ``[0, log(20 Ne)]`` onto 0..255 (0..15); :func:`dequantize`
        inverts them. ``core_block_sites <= 0`` decodes the full sequence."""
        if bits not in (4, 8):
            raise ValueError("bits must be 4 or 8")
        from .device import DeviceArray, run_fb_device

        pi, pj = _split_pairs(pairs)
        P, S = pi.shape[0], self._S
        rows = S if bits == 8 else (S + 1) // 2
        core = int(core_block_sites) if core_block_sites > 0 else max(S, 1)
        flank = int(flank_sites) if core_block_sites > 0 else 0
        codes = np.zeros((rows, P), np.uint8)
        if P and S:
            check_pairs_host(pi, pj, self._n)
            d_pi, d_pj = DeviceArray.from_host(pi), DeviceArray.from_host(pj)
            d_q = DeviceArray((rows, P), np.uint8)
            run_fb_device(self, d_pi, d_pj, P, core, flank, GSMC_Q8 if bits == 8 else GSMC_Q4, d_q)
            d_q.to_host(codes)
        return {"codes": codes, "bits": bits, "log_min": 0.0, "log_max": float(np.log(np.float32(20.0) * self._Ne)),
                "blocks": block_windows(S, core, flank)}

    def run_fwd(self, pairs, mean_only=True):
        """Forward-only filtering over the full sequence
        (``bindings.cpp:2794-2849``): ``mean`` and, unless ``mean_only``,
        ``lower`` / ``upper``, site-major ``(n_sites, n_pairs)`` float32."""
        pi, pj = _split_pairs(pairs)
        P = pi.shape[0]
        out = {"mean": np.zeros((self._S, P), np.float32)}
        if not mean_only:
            out["lower"] = np.zeros((self._S, P), np.float32)
            out["upper"] = np.zeros((self._S, P), np.float32)
        if P and self._S:
            check(lib().gsmc_run_fwd(self._h, _ptr(pi), _ptr(pj), P, 0 if mean_only else GSMC_CI, _ptr(out["mean"]),
                                     _ptr(out.get("lower")), _ptr(out.get("upper"))))
        return out

    def run_fwd_states(self, pairs):
        """Forward states of the cached forward-backward walk over the full
        sequence (``bindings.cpp:2851-2893``): ``mean_log10`` and
        ``cv_log10``, site-major ``(n_sites, n_pairs)`` float32, the state
        (log10 mean in scaled time, log10 CV) after each site."""
        pi, pj = _split_pairs(pairs)
        P = pi.shape[0]
        out = {"mean_log10": np.zeros((self._S, P), np.float32), "cv_log10": np.zeros((self._S, P), np.float32)}
        if P and self._S:
            check(lib().gsmc_run_fwd_states(self._h, _ptr(pi), _ptr(pj), P, _ptr(out["mean_log10"]),
                                            _ptr(out["cv_log10"])))
        return out

    def run_fb_summary(self, pairs, core_block_sites=0, flank_sites=0):
        """Per-site mean, min and max of the posterior mean over ``pairs``,
        without per-pair outputs (``bindings.cpp:2695-2788``): returns
        ``site_mean``, ``site_min``, ``site_max`` (float32, one per site) and
        ``n_pairs``; no pairs gives an empty ``site_mean`` (and ``n_pairs`` 0). The reference decodes the full sequence;
        ``core_block_sites > 0`` decodes block windows instead."""
        pi, pj = _split_pairs(pairs)
        P = pi.shape[0]
        if P == 0:
            return {"site_mean": np.zeros(0, np.float32), "n_pairs": 0}
        if flank_sites < 0:
            raise ValueError("flank_sites must be non-negative.")
        out = {k: np.zeros(self._S, np.float32) for k in ("site_mean", "site_min", "site_max")}
        if self._S:
            check(lib().gsmc_run_fb_summary(self._h, _ptr(pi), _ptr(pj), P, int(core_block_sites), int(flank_sites),
                                            _ptr(out["site_mean"]), _ptr(out["site_min"]), _ptr(out["site_max"])))
        out["n_pairs"] = P
        return out

    def run_fb_gene_stats(self, pairs, gene_sites, core_block_sites=0, flank_sites=0, bin_edges=None,
                          accumulators=None, pair_batch_size=-1):
        """Per-gene statistics of the posterior means over ``pairs``, without
        per-pair outputs: the accumulators of the reference's per-gene
        pipeline (``analysis/genome_wide/infer_chromosome.py:157-227``) for
        genes given as site ranges ``gene_sites[g] = (lo, hi)`` (sites
        ``lo <= s < hi`` of this context). Returns ``count``, ``lin_sum``,
        ``log_sum``, ``log_sq_sum``, ``min_lin``, ``min_log``, ``histogram``
        and ``bin_edges``; pass the result back as ``accumulators`` to add
        further pairs to it. ``core_block_sites <= 0`` decodes the full
        sequence; ``pair_batch_size`` caps the pairs per launch (-1: as many
        as device memory holds)."""
        from .genes import DEFAULT_BIN_EDGES, new_accumulators

        gs = np.ascontiguousarray(np.asarray(gene_sites, dtype=np.int64).reshape(-1, 2))
        edges = np.ascontiguousarray(DEFAULT_BIN_EDGES if bin_edges is None else bin_edges, dtype=np.float64)
        if edges.ndim != 1 or edges.size < 2:
            raise ValueError("bin_edges must hold at least two edges")
        acc = new_accumulators(gs.shape[0], edges) if accumulators is None else accumulators
        want = {"count": (np.int64, (gs.shape[0],)), "lin_sum": (np.float64, (gs.shape[0],)),
                "log_sum": (np.float64, (gs.shape[0],)), "log_sq_sum": (np.float64, (gs.shape[0],)),
                "min_lin": (np.float64, (gs.shape[0],)), "min_log": (np.float64, (gs.shape[0],)),
                "histogram": (np.int64, (gs.shape[0], edges.size - 1))}
        for k, (dt, shape) in want.items():
            a = acc.get(k)
            # the library adds into these arrays in place
            if not (isinstance(a, np.ndarray) and a.dtype == dt and a.shape == shape and a.flags.c_contiguous
                    and a.flags.writeable):
                raise ValueError(f"accumulators[{k!r}] must be a writeable C-contiguous {np.dtype(dt).name} "
                                 f"array of shape {shape} (the genes and bins of this call)")
        if flank_sites < 0:
            raise ValueError("flank_sites must be non-negative.")
        pi, pj = _split_pairs(pairs)
        lo, hi = np.ascontiguousarray(gs[:, 0]), np.ascontiguousarray(gs[:, 1])
        if pair_batch_size == 0 or pair_batch_size < -1:
            raise ValueError("pair_batch_size must be positive or -1 for auto.")
        check(lib().gsmc_run_fb_gene_stats(self._h, _ptr(pi), _ptr(pj), pi.shape[0], int(core_block_sites),
                                           int(flank_sites), int(pair_batch_size), gs.shape[0], _ptr(lo), _ptr(hi),
                                           edges.size - 1,
                                           _ptr(edges), _ptr(acc["count"]), _ptr(acc["lin_sum"]),
                                           _ptr(acc["log_sum"]), _ptr(acc["log_sq_sum"]), _ptr(acc["min_lin"]),
                                           _ptr(acc["min_log"]), _ptr(acc["histogram"])))
        return acc

    def checkpoint_bytes(self, n_sites, core_block_sites, flank_sites, n_pairs):
        out = ctypes.c_uint64()
        check(lib().gsmc_checkpoint_bytes(self._h, int(n_sites), int(core_block_sites), int(flank_sites),
                                          int(n_pairs), ctypes.byref(out)))
        return int(out.value)

    def _alloc(self, rows, cols, mean_only, return_posterior):
        out = {"mean": np.empty((rows, cols), np.float32)}
        if not mean_only:
            out["lower"] = np.empty((rows, cols), np.float32)
            out["upper"] = np.empty((rows, cols), np.float32)
        if return_posterior:
            out["posterior_alpha"] = np.empty((rows, cols), np.float32)
            out["posterior_beta"] = np.empty((rows, cols), np.float32)
        return out

    @staticmethod
    def _flags(mean_only, return_posterior):
        return (0 if mean_only else GSMC_CI) | (GSMC_POSTERIOR if return_posterior else 0)

    def run_fb(self, pairs, mean_only=True, return_posterior=False):
        """Full-sequence forward-backward (``bindings.cpp:2897-2989``).

        Returns ``{"mean": (S, P) float32}`` plus ``lower``/``upper`` when
        ``mean_only`` is false and ``posterior_alpha``/``posterior_beta``
        when ``return_posterior``.
        """
        pi, pj = _split_pairs(pairs)
        P = pi.shape[0]
        if P == 0:
            out = {"mean": np.zeros((self._S, 0), np.float32)}
            if return_posterior:
                out["posterior_alpha"] = np.zeros((self._S, 0), np.float32)
                out["posterior_beta"] = np.zeros((self._S, 0), np.float32)
            return out
        out = self._alloc(self._S, P, mean_only, return_posterior)
        check(lib().gsmc_run_fb(self._h, _ptr(pi), _ptr(pj), P, self._flags(mean_only, return_posterior),
                                _ptr(out["mean"]), _ptr(out.get("lower")), _ptr(out.get("upper")),
                                _ptr(out.get("posterior_alpha")), _ptr(out.get("posterior_beta"))))
        return out

    def run_fb_blockwise(self, pairs, core_block_sites=8192, flank_sites=2048, pair_batch_size=256,
                         max_streams=1, mean_only=True, return_posterior=False):
        """Blockwise forward-backward (``bindings.cpp:2991-3345``); adds ``blocks``."""
        if core_block_sites <= 0:
            raise ValueError("core_block_sites must be positive.")
        if flank_sites < 0:
            raise ValueError("flank_sites must be non-negative.")
        if pair_batch_size == 0 or pair_batch_size < -1:
            raise ValueError("pair_batch_size must be positive or -1 for auto.")
        if max_streams <= 0:
            raise ValueError("max_streams must be positive.")
        if return_posterior and max_streams > 1:
            raise ValueError("return_posterior=True is currently only supported with max_streams=1. "
                             "Run blockwise posterior decoding in single-stream mode.")
        pi, pj = _split_pairs(pairs)
        P = pi.shape[0]
        blocks = block_windows(self._S, core_block_sites, flank_sites)
        if P == 0:
            out = {"mean": np.zeros((self._S, 0), np.float32)}
            if return_posterior:
                out["posterior_alpha"] = np.zeros((self._S, 0), np.float32)
                out["posterior_beta"] = np.zeros((self._S, 0), np.float32)
            out["blocks"] = blocks
            return out
        out = self._alloc(self._S, P, mean_only, return_posterior)
        if self._S:
            check(lib().gsmc_run_fb_blockwise(self._h, _ptr(pi), _ptr(pj), P, int(core_block_sites),
                                              int(flank_sites), int(pair_batch_size), int(max_streams),
                                              self._flags(mean_only, return_posterior), _ptr(out["mean"]),
                                              _ptr(out.get("lower")), _ptr(out.get("upper")),
                                              _ptr(out.get("posterior_alpha")), _ptr(out.get("posterior_beta"))))
        out["blocks"] = blocks
        return out


def gamma_smc_flow_cached_fb(G, positions, pairs, Ne=10000.0, mu=1.25e-8, rho=1e-8, flow_field_path=DEFAULT_PATH,
                             mean_only=False, cache_steps=0):
    """One-shot cached forward-backward over the full sequence
    (``bindings.cpp:2303-2440``, binding ``:3618-3628``): a context for
    ``(G, positions, Ne, mu, rho)`` -- the given rates, no heterozygosity
    estimate -- decodes ``pairs`` and is released. Returns ``mean`` and,
    unless ``mean_only``, ``lower`` / ``upper``, site-major float32."""
    ctx = FlowContext(G, positions, Ne, mu, rho, flow_field_path, cache_steps)
    try:
        return ctx.run_fb(pairs, mean_only=mean_only)
    finally:
        ctx.close()


def gamma_smc_flow_cached_fwd(G, positions, pairs, Ne=10000.0, mu=1.25e-8, rho=1e-8, flow_field_path=DEFAULT_PATH,
                              mean_only=True, cache_steps=0):
    """One-shot forward-only filter (binding ``bindings.cpp:3607-3616``):
    ``FlowContext.run_fwd`` on a context released afterwards."""
    ctx = FlowContext(G, positions, Ne, mu, rho, flow_field_path, cache_steps)
    try:
        return ctx.run_fwd(pairs, mean_only=mean_only)
    finally:
        ctx.close()
