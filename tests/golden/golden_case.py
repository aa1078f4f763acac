"""Committed golden vectors for the decoder path.

``make()`` writes ``golden_case.npz``: a 6-haplotype x 300-site synthetic
cohort (seed 21, 1/k spectrum, positions on [0, 200 kb)), its 16 pairs
(all 15 plus a self pair), the auto-estimated rates, and the FP32 oracle's
full-sequence outputs (mean, lower, upper, posterior alpha/beta) and
blockwise mean (core 96, flank 40). ``make_digests()`` writes the SHA-256 of
every table the reference's own ``build_flow_field_cache`` (oracle/_ref)
produces for two parameter sets. Run ``python -m tests.golden.golden_case``
from the repository root to regenerate both.
"""

import hashlib
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
NPZ = os.path.join(HERE, "golden_case.npz")
DIGESTS = os.path.join(HERE, "ref_table_digests.json")
FF = os.path.join(ROOT, "gamma_smc_cu_amd", "default_flow_field.txt")
CORE, FLANK = 96, 40


def _inputs():
    from gamma_smc_cu_amd import synthetic
    from oracle import oracle as O

    G, pos = synthetic.cohort(6, 300, 200_000, 21)
    pairs = np.array(O.all_pairs(6) + [(2, 2)], np.int32)
    mu, rho = O.estimate_scaled_params(G, pos, 1.25e-8, 1e-8, 1e4)
    return G, pos, pairs, mu, rho


def compute_with_oracle(case):
    from oracle import oracle as O

    G, pos = case["G"], case["positions"]
    pairs = [tuple(map(int, p)) for p in case["pairs"]]
    mu, rho = float(case["mu"]), float(case["rho"])
    T = O.tables_f32(FF, 1000, 1e4, mu, rho)
    r = O.fb_window_f32(G, pos, pairs, T, 1e4, ci=True, posterior=True)
    b = O.fb_blockwise_f32(G, pos, pairs, T, 1e4, CORE, FLANK)
    return {"mean": r["mean"], "lower": r["lower"], "upper": r["upper"], "posterior_alpha": r["posterior_alpha"],
            "posterior_beta": r["posterior_beta"], "blockwise_mean": b["mean"], "blocks": b["blocks"]}


def make():
    G, pos, pairs, mu, rho = _inputs()
    case = {"G": G, "positions": pos, "pairs": pairs, "mu": np.float64(mu), "rho": np.float64(rho)}
    case.update(compute_with_oracle(case))
    np.savez_compressed(NPZ, **case)


def load():
    with np.load(NPZ, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def make_digests():
    from oracle import oracle as O

    out = {}
    for N, Ne, mu, rho in [(1000, 1e4, 1.25e-8, 1e-8), (37, 5000.0, 2e-8, 5e-8)]:
        t = O.ref_tables(FF, N, Ne, mu, rho)
        out[f"N={N},Ne={Ne},mu={mu},rho={rho}"] = {
            name: hashlib.sha256(np.ascontiguousarray(t[k]).tobytes()).hexdigest()
            for k, name in enumerate(O.TABLE_NAMES)}
    with open(DIGESTS, "w") as fh:
        json.dump(out, fh, indent=1, sort_keys=True)


if __name__ == "__main__":
    import sys

    sys.path.insert(0, ROOT)
    make()
    make_digests()
    print("wrote", NPZ, DIGESTS)
