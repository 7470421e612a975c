"""The benchmark configurations of the correction path on the GPU, each at its own size.

  CPU-runnable: 64^3 Gaussian mixture (8 bumps, 0.01 noise), eps = 1e-3 range, separatrix
      geometry, seeds 0..4: bit-exact with the oracle.
  scale: 512^3 fBm (P(k) ~ k^-3 on [0, 1]) at eps = 1e-2, 1e-3 and 1e-4 of the range, too large
      for the oracle: zero mismatches at the end, every value within eps, and the classifier
      finds exactly f's critical cells in the corrected field.
  batched: the 48 fields of the batched benchmark (sides 32..192, Gaussian mixtures and fBm,
      eps log-uniform in [1e-4, 1e-2] range) through one dmtz_correct_batch call: the same
      checks per field, and the three smallest fields bit-exact with the oracle.
  learned: the committed network on 384^3 fields with fresh seeds: the warm-started correction
      reaches zero mismatches within eps with f's critical cells.
"""
import os
import sys

import numpy as np
import pytest

import dmtz_amd
from dmtz_amd import fields

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _critical_cells(ctx, x):
    from test_gpu_parity import _critical_cells as cells

    counts, lo, hi = dmtz_amd.classify(x, ctx=ctx, masks=True)
    return counts, cells(lo, hi, x.shape)


def _same_critical_cells(ctx, f, g):
    cf, a = _critical_cells(ctx, f)
    cg, b = _critical_cells(ctx, g)
    return cf == cg and all(np.array_equal(p, q) for p, q in zip(a, b))


@pytest.mark.parametrize("seed", range(5))
def test_cpu_workload_all_seeds_bit_exact(oracle, dev, seed):
    """Seeds 2, 3 and 4 reach a fixed point with mismatches left, where the reference loops
    forever: both sides report the stall in the same state, and with the finisher both reach
    zero mismatches in the same state."""
    f = fields.gaussian_mixture((64, 64, 64), 8, seed=seed)
    eps, rng = fields.eps_from_rel(f, 1e-3)
    g = fields.quantize(f, eps)
    for fl in (dict(preserve_vpath=1, preserve_geometry=1), dict(preserve_vpath=1, preserve_geometry=1, finish=1)):
        print("64^3 seed", seed, fl, flush=True)
        want = oracle.correct(f, g, eps, rng, oracle.flags(**fl))
        got = dmtz_amd.correct(f, g, eps, rng, dmtz_amd.CorrectionFlags(**fl), ctx=dev, return_counter=True)
        assert got.status == want["status"] and got.status in (0, 2)
        assert got.sweeps == want["sweeps"] and got.counts == want["counts"]
        assert (got.trajectory == want["traj"]).all()
        assert (got.counter == want["counter"]).all()
        assert got.g.tobytes() == want["g"].tobytes()
        if fl.get("finish"):
            assert got.status == 0 and got.counts == (0,) * 8


@pytest.mark.timeout(600)
@pytest.mark.parametrize("rel", [1e-2, 1e-4])
def test_fbm_512_other_bounds(dev, rel):
    """The 512^3 fBm workload at the other two bounds of the scale configuration (1e-3 is
    test_gpu_parity.py::test_fbm_512_converges_with_f_critical_cells)."""
    f = fields.fbm((512, 512, 512), seed=0)
    eps, rng = fields.eps_from_rel(f, rel)
    g = fields.quantize(f, eps)
    fl = dmtz_amd.CorrectionFlags(preserve_vpath=1, preserve_geometry=1, finish=1, max_sweeps=2000)
    print("fbm 512^3 correction, eps/range", rel, flush=True)
    r = dmtz_amd.correct(f, g, eps, rng, fl, ctx=dev)
    assert r.status == 0 and r.counts == (0,) * 8
    assert np.abs(r.g.astype(np.float64) - f).max() <= eps
    assert _same_critical_cells(dev, f, r.g)


def _batch_fields(k, seed):
    sys.path.insert(0, ROOT)
    import bench

    return bench.batch_fields(k, seed)


@pytest.mark.timeout(600)
def test_batched_48_fields(oracle, dev):
    items = _batch_fields(48, 0)
    fl = dmtz_amd.CorrectionFlags(preserve_vpath=1, preserve_geometry=1, finish=1, max_sweeps=5000)
    print("batch of %d fields, %d voxels" % (len(items), sum(f.size for f, *_ in items)), flush=True)
    res = dmtz_amd.correct_batch([(f, g, eps, rng) for f, g, eps, rng, _ in items], fl, ctx=dev)
    for i, ((f, g, eps, rng, rel), r) in enumerate(zip(items, res)):
        assert r.status == 0 and r.counts == (0,) * 8, (i, f.shape, r.status, r.counts)
        assert np.abs(r.g.astype(np.float64) - f).max() <= eps, i
        assert _same_critical_cells(dev, f, r.g), (i, f.shape)
    order = np.argsort([f.size for f, *_ in items])[:3]
    ofl = oracle.flags(preserve_vpath=1, preserve_geometry=1, finish=1, max_sweeps=5000)
    for i in order.tolist():
        f, g, eps, rng, _ = items[i]
        print("oracle on batch field", i, f.shape, flush=True)
        want = oracle.correct(f, g, eps, rng, ofl)
        assert res[i].sweeps == want["sweeps"] and res[i].g.tobytes() == want["g"].tobytes(), i


@pytest.mark.timeout(600)
def test_learned_on_384_held_out_fields(dev):
    from dmtz_amd import devmem
    from dmtz_amd import learned as L

    net = L.Net(seed=0, ctx=dev)
    net.set_params(np.load(os.path.join(ROOT, "dmtz_amd", "net_weights.npy")))
    fl = dmtz_amd.CorrectionFlags(preserve_vpath=1, preserve_geometry=1, finish=1, max_sweeps=5000)
    for kind, seed, rel in [("fbm", 901, 3e-3), ("gm", 902, 3e-3)]:
        if kind == "fbm":
            f = fields.fbm((384, 384, 384), seed=seed)
        else:
            f = fields.gaussian_mixture((384, 384, 384), 16, seed=seed, noise=0.01)
        eps, rng = fields.eps_from_rel(f, rel)
        g0 = fields.quantize(f, eps)
        print("learned 384^3", kind, flush=True)
        fd = devmem.DeviceArray.from_numpy(f)
        gd = devmem.DeviceArray.from_numpy(g0)
        r, warm = L.correct_learned(net, fd, gd, eps, rng, fl)
        g1 = gd.numpy()
        assert r.status == 0 and r.counts == (0,) * 8
        assert np.abs(g1.astype(np.float64) - f).max() <= eps
        assert _same_critical_cells(dev, f, g1)
    net.close()


def test_batched_learned_small(dev):
    """dmtz_correct_batch_learned_f32 on a few uneven fields: every field converges within eps
    with f's critical cells, and the plain loop decides the result as in the single-field form."""
    from dmtz_amd import devmem
    from dmtz_amd import learned as L

    net = L.Net(seed=0, ctx=dev)
    net.set_params(np.load(os.path.join(ROOT, "dmtz_amd", "net_weights.npy")))
    fl = dmtz_amd.CorrectionFlags(preserve_vpath=1, preserve_geometry=1, finish=1)
    items, host = [], []
    for seed, shape, rel in [(0, (40, 33, 27), 1e-2), (1, (24, 48, 31), 3e-3), (2, (50, 20, 36), 1e-3)]:
        f = fields.gaussian_mixture(shape, 6, seed=seed, noise=0.01)
        eps, rng = fields.eps_from_rel(f, rel)
        g0 = fields.quantize(f, eps)
        host.append((f, g0, eps))
        items.append((devmem.DeviceArray.from_numpy(f), devmem.DeviceArray.from_numpy(g0), eps, rng))
    res, warm = L.correct_batch_learned(net, items, fl)
    for (f, g0, eps), (fd, gd, _, rng), r, w in zip(host, items, res, warm):
        g1 = gd.numpy()
        assert r.status == 0 and r.counts == (0,) * 8
        assert np.abs(g1.astype(np.float64) - f).max() <= eps
        assert _same_critical_cells(dev, f, g1)
        assert w >= 0
        one_g = devmem.DeviceArray.from_numpy(g0)
        one, w1 = L.correct_learned(net, fd, one_g, eps, rng, fl)
        assert w1 == w and one.sweeps == r.sweeps and one_g.numpy().tobytes() == g1.tobytes()
    net.close()
