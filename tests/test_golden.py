"""Golden fixtures: tests/golden/fixtures.json, written from the oracle by
tests/golden/make_golden.py (the reference ships no tests or vectors of its own).

CPU: the oracle still reproduces the small fixtures (a regression pin of the oracle itself).
GPU: the HIP path reproduces every fixture bit for bit through the C ABI, including the full-size
256^3 benchmark workload in float64 (40 sweeps to the reference's fixed point) and, when present,
in float32 with the finisher: status, sweeps, counts, the per-sweep trajectory, and SHA-256 hashes
of the corrected field and of the delta counters.
"""
import hashlib
import importlib.util
import json
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
_spec = importlib.util.spec_from_file_location("make_golden", os.path.join(HERE, "golden", "make_golden.py"))
MG = importlib.util.module_from_spec(_spec)
_spec.loader.exec_module(MG)
FX = json.load(open(os.path.join(HERE, "golden", "fixtures.json")))
SMALL = sorted(k for k, c in FX.items() if np.prod(c["shape"]) <= 64 ** 3)


def _sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def _check(c, status, sweeps, counts, traj, g, counter):
    assert status == c["status"]
    assert sweeps == c["sweeps"]
    assert list(counts) == c["counts"]
    assert np.array_equal(np.asarray(traj), np.asarray(c["trajectory"], dtype=np.int64).reshape(-1, 8))
    assert _sha(g) == c["g_sha256"]
    assert _sha(counter.astype(np.uint8)) == c["counter_sha256"]


@pytest.mark.parametrize("name", SMALL)
def test_oracle_reproduces_fixture(oracle, name):
    c = FX[name]
    f, g, eps, rng = MG.case_input(c)
    r = oracle.correct(f, g, eps, rng, oracle.flags(**MG.FLAGS[c["flags"]], max_sweeps=5000), traj_cap=8192)
    _check(c, r["status"], r["sweeps"], r["counts"], r["traj"], r["g"], r["counter"])


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(FX))
def test_device_reproduces_fixture(dev, name):
    import dmtz_amd

    c = FX[name]
    print("fixture", name, flush=True)
    f, g, eps, rng = MG.case_input(c)
    fl = dmtz_amd.CorrectionFlags(**MG.FLAGS[c["flags"]], max_sweeps=5000)
    r = dmtz_amd.correct(f, g, eps, rng, fl, ctx=dev, return_counter=True)
    _check(c, r.status, r.sweeps, r.counts, r.trajectory, r.g, r.counter)
    if r.status == 0:
        assert np.abs(r.g.astype(np.float64) - f).max() <= eps
