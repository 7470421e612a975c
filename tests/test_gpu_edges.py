"""Edge cases of the correction on the GPU, bit-exact with the oracle where it can run.

  - a constant field (every value tied), and g equal to f: nothing to repair, g unchanged after
    the one separatrix check;
  - the smallest grid the path accepts (nx = 3, ny = nz = 2) and thin slabs, random values, in
    every flag mode of the parity corpus;
  - integer fields with heavy ties in both f and g;
  - a ragged 84 M-vertex grid (40 x 700 x 3001), too large for the oracle: zero mismatches, every
    value within eps, and f's critical cells in the corrected field;
  - a grid of more than 2048 x 64 detection tiles of 2048 vertices (268.7 M vertices), whose one
    planted mismatch lies in a tile that the detection grid reaches in its second group of 64
    tiles per block.
"""
import numpy as np
import pytest

import corpus
import dmtz_amd
from dmtz_amd import fields

pytestmark = pytest.mark.gpu


def _compare_run(oracle, dev, f, g, eps, rng, fl):
    want = oracle.correct(f, g, eps, rng, oracle.flags(**fl))
    got = dmtz_amd.correct(f, g, eps, rng, dmtz_amd.CorrectionFlags(**fl), ctx=dev, return_counter=True)
    assert got.status == want["status"]
    assert got.sweeps == want["sweeps"] and got.counts == want["counts"]
    assert (got.trajectory == want["traj"]).all()
    assert (got.counter == want["counter"]).all()
    assert got.g.tobytes() == want["g"].tobytes()
    return got


def test_nothing_to_repair(oracle, dev):
    fl = dict(preserve_vpath=1, preserve_geometry=1, finish=1)
    f = np.full((5, 7, 9), 0.5, np.float32)
    for g, eps in [(f.copy(), 0.1), (f - np.float32(0.05), 0.1)]:  # all ties; a uniform shift keeps every tie
        got = _compare_run(oracle, dev, f, g, eps, 1.0, fl)
        assert got.status == 0 and got.sweeps == 1 and got.counts == (0,) * 8  # the one separatrix check
        assert got.g.tobytes() == g.tobytes()
    f = fields.gaussian_mixture((12, 15, 17), 5, seed=3, noise=0.02)
    eps, rng = fields.eps_from_rel(f, 1e-2)
    got = _compare_run(oracle, dev, f, f.copy(), eps, rng, fl)
    assert got.sweeps == 1 and got.g.tobytes() == f.tobytes()


def _small_grids():
    out = []
    r = np.random.default_rng(2024)
    for shape in [(2, 2, 3), (2, 3, 3), (3, 2, 3), (2, 2, 9), (2, 11, 3), (13, 2, 3), (2, 40, 90), (40, 2, 90)]:
        for k in range(2):
            f = r.standard_normal(shape).astype(np.float32)
            eps = 0.3
            g = (f + r.uniform(-eps, eps, shape)).astype(np.float32)
            g = np.clip(g, f - np.float32(eps), f + np.float32(eps))
            out.append(("rand%s_%d" % ("x".join(map(str, shape)), k), f, g, eps, float(f.max() - f.min())))
    for k in range(3):  # heavy ties in f and in g
        shape = (6, 7, 8)
        f = r.integers(0, 3, size=shape).astype(np.float32)
        g = (f - r.integers(0, 2, size=shape)).astype(np.float32)
        out.append(("ties_%d" % k, f, g, 1.0, float(f.max() - f.min())))
    return out


@pytest.mark.parametrize("name,fl", corpus.FLAG_CASES, ids=[c[0] for c in corpus.FLAG_CASES])
def test_small_and_thin_grids_bit_exact(oracle, dev, name, fl):
    for inp, f, g, eps, rng in _small_grids():
        print("edge", name, inp, f.shape, flush=True)
        got = _compare_run(oracle, dev, f, g, eps, rng, fl)
        if got.status == 0:
            assert got.counts == (0,) * 8
            assert np.abs(got.g.astype(np.float64) - f).max() <= eps


@pytest.mark.timeout(600)
def test_ragged_84m_grid(dev):
    from test_gpu_parity import _critical_cells

    shape = (40, 700, 3001)
    f = fields.gaussian_mixture(shape, 32, seed=5, noise=0.01)
    eps, rng = fields.eps_from_rel(f, 3e-3)
    g = fields.quantize(f, eps)
    fl = dmtz_amd.CorrectionFlags(preserve_vpath=1, preserve_geometry=1, finish=1)
    print("ragged", shape, flush=True)
    r = dmtz_amd.correct(f, g, eps, rng, fl, ctx=dev)
    assert r.status == 0 and r.counts == (0,) * 8
    assert np.abs(r.g.astype(np.float64) - f).max() <= eps
    cf, lof, hif = dmtz_amd.classify(f, ctx=dev, masks=True)
    cg, log_, hig = dmtz_amd.classify(r.g, ctx=dev, masks=True)
    assert cf == cg
    for a, b in zip(_critical_cells(lof, hif, shape), _critical_cells(log_, hig, shape)):
        assert np.array_equal(a, b)


@pytest.mark.timeout(600)
def test_detection_past_the_first_group_of_tiles(dev):
    shape = (65600, 64, 64)  # nz, ny, nx: 268,697,600 vertices, 131,200 tiles > 2048 * 64
    n = int(np.prod(shape))
    assert n > 2048 * 64 * 2048
    # f increases with the vertex index (float32 ties are broken by index): no critical cell but
    # the minimum. g lowers one vertex of the last tiles far below its neighbours.
    f = np.arange(n, dtype=np.int64).astype(np.float32).reshape(shape)
    g = f.copy()
    z, y, x = shape[0] - 2, 30, 30
    v0 = (z * shape[1] + y) * shape[2] + x
    assert v0 // 2048 >= 2048 * 64
    eps = 1e5
    g[z, y, x] = f[z, y, x] - np.float32(5e4)
    fl = dmtz_amd.CorrectionFlags(preserve_vpath=0, preserve_geometry=0, finish=1)
    print("detection past the first group", shape, flush=True)
    r = dmtz_amd.correct(f, g, eps, float(n - 1), fl, ctx=dev)
    assert r.status == 0 and r.counts == (0,) * 8
    assert r.sweeps >= 2 and r.trajectory[0].any()  # the planted mismatch was found
    assert np.abs(r.g.astype(np.float64) - f).max() <= eps
    assert dmtz_amd.classify(r.g, ctx=dev) == dmtz_amd.classify(f, ctx=dev)
