"""GPU test of the entry point that allocates nothing
(``gsmc_run_fb_device_noalloc``): called through the C ABI, its outputs
equal ``gsmc_run_fb_device``'s bit for bit, on two calls in a row (the
second smaller, reusing the device scratch the first one wrote), with and
without CI / posterior outputs; a call beyond the scratch maximum is
rejected before anything runs, and the other entry points still work."""

import ctypes

import numpy as np
import pytest

import gamma_smc_cu_amd
from gamma_smc_cu_amd import _core, synthetic
from gamma_smc_cu_amd.device import DeviceArray, Stream, run_fb_device
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _case(msg):
    print(f"[case] {msg}", flush=True)


def _noalloc(ctx, d_pi, d_pj, P, core, flank, flags, outs, stream):
    p = [None if o is None else o.ptr for o in outs]
    _core.check(_core.lib().gsmc_run_fb_device_noalloc(ctx.handle, d_pi.ptr, d_pj.ptr, int(P), int(core), int(flank),
                                                       int(flags), *p, stream.handle))


def _outs(S, P, flags):
    ci, post = bool(flags & _core.GSMC_CI), bool(flags & _core.GSMC_POSTERIOR)
    return [DeviceArray((S, P), np.float32), DeviceArray((S, P), np.float32) if ci else None,
            DeviceArray((S, P), np.float32) if ci else None, DeviceArray((S, P), np.float32) if post else None,
            DeviceArray((S, P), np.float32) if post else None]


@pytest.fixture(scope="module")
def cohort(flow_field_path):
    G, pos = synthetic.cohort(20, 3000, 900_000, 77)
    mu, rho = O.estimate_scaled_params(G, pos, 1.25e-8, 1e-8, 1e4)
    ctx = gamma_smc_cu_amd.FlowContext(G, pos, 1e4, mu, rho, flow_field_path, 0)
    return G, pos, ctx


@pytest.mark.parametrize("flags", [0, _core.GSMC_CI | _core.GSMC_POSTERIOR])
def test_noalloc_matches_device_entry_twice(cohort, flags):
    G, pos, ctx = cohort
    S = G.shape[1]
    st = Stream()
    allp = np.array(O.all_pairs(G.shape[0]), np.int32)
    for P, core, flank in ((190, 1000, 150), (37, 512, 64)):  # second call smaller
        _case(f"noalloc flags={flags} P={P} core={core} flank={flank}")
        pairs = allp[:P]
        d_pi = DeviceArray.from_host(np.ascontiguousarray(pairs[:, 0]))
        d_pj = DeviceArray.from_host(np.ascontiguousarray(pairs[:, 1]))
        ref, got = _outs(S, P, flags), _outs(S, P, flags)
        run_fb_device(ctx, d_pi, d_pj, P, core, flank, flags, *ref, stream=st)
        _noalloc(ctx, d_pi, d_pj, P, core, flank, flags, got, st)
        st.synchronize()
        for r, g in zip(ref, got):
            if r is not None:
                assert np.array_equal(r.to_host().view(np.uint32), g.to_host().view(np.uint32))


def test_noalloc_rejects_calls_beyond_its_scratch(cohort):
    G, pos, ctx = cohort
    S = G.shape[1]
    P = 64 * 400  # 400 tiles x (3000 / 8 + 2) checkpoint rows x 512 B > 64 MiB
    d_pi = DeviceArray.from_host(np.zeros(P, np.int32))
    d_pj = DeviceArray.from_host(np.ones(P, np.int32))
    dummy = DeviceArray((1,), np.float32)
    st = Stream()
    _case("noalloc oversize")
    with pytest.raises(ValueError, match="scratch"):
        _core.check(_core.lib().gsmc_run_fb_device_noalloc(ctx.handle, d_pi.ptr, d_pj.ptr, P, S, 0, 0, dummy.ptr,
                                                           None, None, None, None, st.handle))
    # nothing ran; the context and the allocating entry point still work
    pairs = [(1, 0), (3, 2)]
    r = ctx.run_fb(pairs)
    assert np.isfinite(r["mean"]).all() and (r["mean"] > 0).all()
