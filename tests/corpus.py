"""The parity corpus: the inputs on which the GPU parity tests compare the HIP path with the
oracle, shared with the CPU test that checks which of the oracle's restated branches these
inputs reach (test_oracle_coverage.py). Every generator is seeded and small.
"""
import numpy as np

from dmtz_amd import fields

# Flag modes of the full-correction comparisons (test_correction_bit_exact and the wall runs).
FLAG_CASES = [
    ("cp", {}),
    ("geometry", dict(preserve_vpath=1, preserve_geometry=1)),
    ("connectivity", dict(preserve_vpath=1, preserve_geometry=0)),
    ("edit_type1", dict(edit_type=1, preserve_vpath=1, preserve_geometry=1)),
    ("filtration", dict(filtration=1)),
    ("preserve_type", dict(preserve_type=1)),
    # without saddle repairs these inputs reach a fixed point with maxima left (the reference
    # loops forever there); both sides must report the stall in the same state
    ("no_saddles", dict(preserve_saddles=0, max_sweeps=300)),
    # the same stalls resolved by the finisher: restored boxes, frozen counters, zero mismatches
    ("no_saddles_finish", dict(preserve_saddles=0, max_sweeps=300, finish=1)),
    # runs that never stall are unchanged by the finisher
    ("geometry_finish", dict(preserve_vpath=1, preserve_geometry=1, finish=1)),
]

WALL_CASES = [
    ("walls_geometry", dict(preserve_vpath=1, preserve_geometry=1, preserve_connectors=1)),
    ("walls_connectivity", dict(preserve_vpath=1, preserve_geometry=0, preserve_connectors=1)),
    ("walls_finish", dict(preserve_vpath=1, preserve_geometry=1, preserve_connectors=1, finish=1)),
]


def gm_inputs():
    """(name, f, g, eps, range): Gaussian mixtures quantised at three bounds."""
    out = []
    for seed, shape, rel in [(0, (16, 16, 16), 1e-2), (1, (24, 20, 12), 2e-2), (2, (13, 17, 11), 5e-2)]:
        f = fields.gaussian_mixture(shape, 6, seed=seed, noise=0.02)
        eps, rng = fields.eps_from_rel(f, rel)
        out.append(("gm%d" % seed, f, fields.quantize(f, eps), eps, rng))
    return out


def fbm_inputs():
    """(name, f, g, eps, range): crops of a 64^3 fBm field (P(k) ~ k^-3, the 512^3 scale
    workload's generator) at the three bounds of that workload."""
    full = fields.fbm((64, 64, 64), seed=7)
    f = np.ascontiguousarray(full[5:29, 9:33, 3:27])  # 24^3
    out = []
    for rel in (1e-2, 1e-3, 1e-4):
        eps, rng = fields.eps_from_rel(f, rel)
        out.append(("fbm%g" % rel, f, fields.quantize(f, eps), eps, rng))
    return out


def _branch_input(seed):
    """Small (f, g) pairs that reach repair and check branches the smooth inputs above do not:
    integer fields with exact snaps (g = f - eps in exact arithmetic), independent random
    orders of f and g, and quantised white noise."""
    r = np.random.default_rng(seed)
    if seed in (543, 965):  # integer values 0..9, eps = 1
        r = np.random.default_rng(10000 + seed)
        shape = tuple(int(x) for x in r.integers(3, 8, size=3))
        shape = (shape[0], shape[1], max(3, shape[2]))
        f = r.integers(0, 10, size=shape).astype(np.float32)
        if seed == 543:
            g = f - r.integers(0, 2, size=shape).astype(np.float32)
        else:
            g = (f + r.integers(-1, 2, size=shape)).astype(np.float32)
        return f, g, 1.0, float(f.max() - f.min())
    if seed == 8106:  # f and g independent random orders of 36 vertices, eps above the range
        r = np.random.default_rng(50000 + seed)
        shape = (4, 3, 3)
        f = r.permutation(36).reshape(shape).astype(np.float32)
        g = r.permutation(36).reshape(shape).astype(np.float32)
        return f, g, 100.0, 100.0
    # seeds 0 and 1407: a quantised Gaussian mixture and quantised white noise
    shape = tuple(int(x) for x in r.integers(4, 9, size=3))
    if seed % 4 == 0:
        f = fields.gaussian_mixture(shape, 3, seed=seed, noise=0.05)
    else:
        f = r.standard_normal(shape).astype(np.float32)
    rel = float(10 ** r.uniform(-2, -0.5))
    eps, rng = fields.eps_from_rel(f, rel)
    if seed % 3 == 0:
        g = fields.quantize(f, eps)
    else:
        g = (f.astype(np.float64) + r.uniform(-eps, eps, size=shape)).astype(np.float32)
        g = np.where(np.abs(g.astype(np.float64) - f) > eps, f, g).astype(np.float32)
    return f, g, eps, rng


def branch_inputs():
    """(name, f, g, eps, range) of the branch-reaching inputs."""
    return [("branch%d" % s,) + _branch_input(s) for s in (0, 1407, 543, 965, 8106)]


# Tiny integer fields of the hand-worked repair-rule cases (test_oracle_kats.py): x fastest,
# 3x3x2 and 3x2x2 grids, (D, H, W) = (2, 3, 3) and (2, 2, 3); g differs at the listed vertices.
F_A = [0, 7, 9, 16, 5, 8, 1, 10, 17, 2, 13, 11, 15, 6, 4, 3, 12, 14]
F_B = [9, 11, 1, 3, 2, 4, 6, 7, 0, 10, 5, 8]
F_C = [8, 11, 4, 7, 5, 0, 1, 9, 2, 10, 6, 3]
F_D = [6, 16, 7, 10, 2, 11, 0, 15, 17, 14, 9, 12, 5, 13, 3, 4, 8, 1]
KAT_INPUTS = {
    "A": ((2, 3, 3), F_A, {8: 0.5}),
    "B": ((2, 2, 3), F_B, {1: 2.5}),
    "C": ((2, 2, 3), F_C, {4: 1.5}),
    "D": ((2, 3, 3), F_D, {2: 11.5, 4: 4.5}),
}


def kat_field(name):
    """(f, g) of the hand-worked repair-rule cases of test_oracle_kats.py."""
    shape, fv, changes = KAT_INPUTS[name]
    f = np.array(fv, np.float32).reshape(shape)
    g = f.copy()
    for v, x in changes.items():
        g.flat[v] = x
    return f, g


def edit_set_inputs():
    """(name, f, g, eps, range, dtype) of the one-phase edit-set comparisons."""
    out = []
    for seed in range(3):
        for shape, rel in [((8, 9, 10), 1e-2), ((16, 16, 16), 3e-2), ((6, 7, 9), 5e-2)]:
            f = fields.gaussian_mixture(shape, 4, seed=seed, noise=0.02)
            eps, rng = fields.eps_from_rel(f, rel)
            out.append(("gm%d_%s" % (seed, shape), f, fields.quantize(f, eps), eps, rng))
    out += fbm_inputs()
    out += branch_inputs()
    out += [("kat_" + k,) + kat_field(k) + (100.0, 100.0) for k in sorted(KAT_INPUTS)]
    return out


def path_inputs(oracle):
    """(name, f, g) of the separatrix-list comparisons: g with critical cells already matching f's
    (the state in which the reference runs these checks) and, for the checks on their own,
    quantised g whose critical cells still differ."""
    out = []
    for seed in range(4):
        for shape, rel in [((12, 12, 12), 1e-2), ((10, 14, 9), 3e-2), ((16, 16, 16), 5e-3)]:
            f = fields.gaussian_mixture(shape, 5, seed=seed, noise=0.02)
            eps, rng = fields.eps_from_rel(f, rel)
            g0 = fields.quantize(f, eps)
            out.append(("gm%d_%s" % (seed, shape), f, oracle.correct(f, g0, eps, rng, oracle.flags())["g"], eps, rng))
    for name, f, g0, eps, rng in fbm_inputs()[:2] + branch_inputs()[:1]:
        out.append((name + "_quantised", f, g0, eps, rng))
        out.append((name, f, oracle.correct(f, g0, eps, rng, oracle.flags())["g"], eps, rng))
    return out


def wall_inputs(oracle):
    """(name, f, g, eps, range) of the wall-list comparisons, quantised and cp-corrected g."""
    out = []
    for seed, shape, rel in [(2, (16, 16, 16), 5e-3), (1, (24, 20, 12), 2e-2), (3, (13, 17, 11), 5e-2)]:
        f = fields.gaussian_mixture(shape, 6, seed=seed, noise=0.02)
        eps, rng = fields.eps_from_rel(f, rel)
        g0 = fields.quantize(f, eps)
        out.append(("gm%d_q" % seed, f, g0, eps, rng))
        out.append(("gm%d_c" % seed, f, oracle.correct(f, g0, eps, rng, oracle.flags())["g"], eps, rng))
    out.append(("branch543",) + _branch_input(543))
    return out
