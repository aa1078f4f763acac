"""The parts of bench.py's contract that need no GPU: the refusal of more
than one GPU, and the algorithmic byte count its roofline divides by."""

import json
import os
import subprocess
import sys

from tests.conftest import ROOT


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    line = json.loads(r.stdout.strip().splitlines()[-1])
    assert "one GPU" in line["error"]


def test_algorithmic_bytes_of_the_chromosome_decode():
    """63,190 pairs x 402,853 sites at checkpoint_k 8: 4 B of posterior mean
    per pair-site (101.8 GB), a 16 B checkpoint record read and written per
    two 8-site windows (2 x 25.5 GB), then the packed rows, descriptors and
    pair ids once: 152.76 GB."""
    import bench

    P, S, n = 63190, 402853, 356
    total = bench.algorithmic_bytes(P, S, n, 8)
    assert abs(P * S * 4 / 1e9 - 101.82) < 0.01
    assert abs(P * S * 2.0 / 1e9 - 50.91) < 0.01
    assert abs(total / 1e9 - 152.76) < 0.01
    # the checkpoint share halves at K = 16 and the stored forward states
    # (K = 1) cost 16 B per pair-site
    assert abs((bench.algorithmic_bytes(P, S, n, 16) - total) / 1e9 + 25.46) < 0.01
    assert abs((bench.algorithmic_bytes(P, S, n, 1) - bench.algorithmic_bytes(P, S, n, 16)) / (P * S) - 15.0) < 1e-9
