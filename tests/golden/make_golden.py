"""Regenerate tests/golden/fixtures.json with the oracle (oracle/dmtz_oracle.cpp).

The reference ships no tests or vectors, so these fixtures are the oracle's results on seeded
inputs: per-sweep mismatch trajectories, final counts and SHA-256 hashes of the corrected field
and of the delta counters. The device path must reproduce them bit for bit. The small cases run in
seconds; the full-size cases (256^3, the benchmark workload) take the oracle tens of minutes on 8
threads and are only regenerated with --full.

    python tests/golden/make_golden.py [--full] [--threads N]
"""
import argparse
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from dmtz_amd import fields  # noqa: E402
from oracle import oracle as O  # noqa: E402

FLAGS = {
    "cp": dict(preserve_saddles=1),
    "geometry": dict(preserve_saddles=1, preserve_vpath=1, preserve_geometry=1),
    "connectivity": dict(preserve_saddles=1, preserve_vpath=1, preserve_geometry=0),
    "edit_type1": dict(preserve_saddles=1, preserve_vpath=1, preserve_geometry=1, edit_type=1),
    "no_saddles_finish": dict(preserve_saddles=0, finish=1),
    "geometry_finish": dict(preserve_saddles=1, preserve_vpath=1, preserve_geometry=1, finish=1),
}

SMALL = [
    dict(name="gm16_f32_geometry", shape=[16, 16, 16], bumps=6, seed=0, noise=0.02, rel=1e-2, dtype="f32", flags="geometry"),
    dict(name="gm24x20x12_f32_cp", shape=[24, 20, 12], bumps=6, seed=1, noise=0.02, rel=2e-2, dtype="f32", flags="cp"),
    dict(name="gm13x17x11_f64_connectivity", shape=[13, 17, 11], bumps=6, seed=2, noise=0.02, rel=5e-2, dtype="f64", flags="connectivity"),
    dict(name="gm20x18x16_f32_edit_type1", shape=[20, 18, 16], bumps=6, seed=4, noise=0.02, rel=2e-2, dtype="f32", flags="edit_type1"),
    dict(name="gm14x10x12_f32_no_saddles_finish", shape=[14, 10, 12], bumps=5, seed=2, noise=0.01, rel=2e-2, dtype="f32", flags="no_saddles_finish"),
    dict(name="gm64_f32_c1_geometry", shape=[64, 64, 64], bumps=8, seed=0, noise=0.01, rel=1e-3, dtype="f32", flags="geometry"),
]
FULL = [
    dict(name="gm256_f64_geometry_stall", shape=[256, 256, 256], bumps=64, seed=0, noise=0.01, rel=1e-2, dtype="f64", flags="geometry"),
    dict(name="gm256_f32_geometry_finish", shape=[256, 256, 256], bumps=64, seed=0, noise=0.01, rel=1e-2, dtype="f32", flags="geometry_finish"),
]


def case_input(c):
    f = fields.gaussian_mixture(tuple(c["shape"]), c["bumps"], seed=c["seed"], noise=c["noise"])
    eps, rng = fields.eps_from_rel(f, c["rel"])
    g = fields.quantize(f, eps)
    dt = np.float64 if c["dtype"] == "f64" else np.float32
    return f.astype(dt), g.astype(dt), eps, rng


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def run(c):
    f, g, eps, rng = case_input(c)
    r = O.correct(f, g, eps, rng, O.flags(**FLAGS[c["flags"]], max_sweeps=5000), traj_cap=8192)
    out = dict(c)
    out.update(status=r["status"], sweeps=r["sweeps"], counts=list(r["counts"]),
               trajectory=r["traj"].tolist(), g_sha256=sha(r["g"]),
               counter_sha256=sha(r["counter"].astype(np.uint8)), stalls=r["stalls"], finished=r["finished"])
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--full", action="store_true")
    ap.add_argument("--threads", type=int, default=8)
    args = ap.parse_args()
    O.build()
    O.set_threads(args.threads)
    path = os.path.join(HERE, "fixtures.json")
    old = json.load(open(path)) if os.path.exists(path) else {}
    keep = {k: v for k, v in old.items() if any(k == c["name"] for c in FULL)} if not args.full else {}
    res = dict(keep)
    for c in SMALL + (FULL if args.full else []):
        print("fixture", c["name"], flush=True)
        res[c["name"]] = run(c)
    with open(path, "w") as fh:
        json.dump(res, fh, indent=0, sort_keys=True)


if __name__ == "__main__":
    main()
