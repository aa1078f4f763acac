"""Shared fixtures. GPU tests are marked ``gpu``; everything else runs on CPU."""

import os
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

FLOW_FIELD = os.path.join(ROOT, "gamma_smc_cu_amd", "default_flow_field.txt")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


def _ensure_built():
    oracle_lib = os.path.join(ROOT, "oracle", "liboracle.so")
    if not os.path.exists(oracle_lib):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "liboracle.so"], check=True)
    lib = os.path.join(ROOT, "gamma_smc_cu_amd", "libgsmc.so")
    if not os.path.exists(lib):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "gamma_smc_cu_amd", "csrc")], check=True)


_ensure_built()


@pytest.fixture(scope="session")
def flow_field_path():
    return FLOW_FIELD


def small_cohort(n=12, S=700, span=400_000, seed=3):
    from gamma_smc_cu_amd import synthetic

    return synthetic.cohort(n, S, span, seed)


@pytest.fixture(scope="session")
def cohort_small():
    return small_cohort()
