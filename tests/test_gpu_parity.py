"""GPU parity of the HIP path against the oracle, through the C ABI (include/dmtz.h).

Integer / index work is bit-exact: pairing arrays, mismatch counts, edit sets, per-sweep count
trajectories, delta counters and the corrected field g' (float bits) must equal the oracle's on
the same seeded inputs.
"""
import numpy as np
import pytest

import corpus
import dmtz_amd
from dmtz_amd import fields

pytestmark = pytest.mark.gpu


def _gradient_cases():
    out = []
    for seed in range(3):
        for shape in [(7, 6, 5), (5, 8, 9), (9, 9, 9), (4, 3, 6)]:
            out.append(fields.gaussian_mixture(shape, 3, seed=seed, noise=0.05))
    rng = np.random.default_rng(1)
    for k in range(3):
        out.append(rng.integers(0, 3, size=(6, 7, 8)).astype(np.float32))
        out.append(rng.integers(0, 2, size=(5, 5, 5)).astype(np.float64))
    out.append(fields.gaussian_mixture((17, 19, 23), 6, seed=9, noise=0.02))
    return out


def test_gradient_bit_exact(oracle, dev):
    for f in _gradient_cases():
        print("gradient", f.shape, f.dtype, flush=True)
        A = oracle.gradient(f)
        B = dmtz_amd.gradient(f, ctx=dev)
        for name, a, b in zip(("vertices", "edges", "faces", "tets"), A, B):
            assert (a == b).all(), (f.shape, name, np.nonzero(a != b)[0][:5])


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("edit_type", [0, 1])
def test_edit_set_bit_exact(oracle, dev, dtype, edit_type):
    """One detection + repair phase: mismatch counts and edit sets, on Gaussian mixtures, fBm
    crops and the inputs that reach the rarer repair branches (tests/corpus.py)."""
    for name, f, g, eps, rng in corpus.edit_set_inputs():
        f = f.astype(dtype)
        g = g.astype(dtype)
        print("targets", name, dtype.__name__, edit_type, flush=True)
        want, wc = oracle.targets(f, g, eps, rng, oracle.flags(edit_type=edit_type))
        got, gc = dmtz_amd.targets(f, g, eps, rng, dmtz_amd.CorrectionFlags(edit_type=edit_type),
                                   ctx=dev)
        assert gc == wc, name
        assert (got == want).all(), name


FLAG_CASES = corpus.FLAG_CASES
# expected status per case (0 converged, 2 stalled)
EXPECT = {"no_saddles": 2}


def _compare_run(oracle, dev, f, g, eps, rng, fl):
    want = oracle.correct(f, g, eps, rng, oracle.flags(**fl))
    got = dmtz_amd.correct(f, g, eps, rng, dmtz_amd.CorrectionFlags(**fl), ctx=dev,
                           return_counter=True)
    assert got.status == want["status"]
    assert got.sweeps == want["sweeps"]
    assert got.counts == want["counts"]
    assert (got.trajectory == want["traj"]).all()
    assert (got.counter == want["counter"]).all()
    assert got.g.tobytes() == want["g"].tobytes()
    return got


@pytest.mark.parametrize("name,fl", FLAG_CASES, ids=[c[0] for c in FLAG_CASES])
def test_correction_bit_exact(oracle, dev, name, fl):
    """Full corrections on Gaussian mixtures and on fBm crops at eps = 1e-2, 1e-3 and 1e-4 of the
    range (tests/corpus.py), compared with the oracle bit for bit."""
    for inp, f, g, eps, rng in corpus.gm_inputs() + corpus.fbm_inputs():
        print("correct", name, inp, f.shape, flush=True)
        got = _compare_run(oracle, dev, f, g, eps, rng, fl)
        if inp.startswith("gm"):
            assert got.status == EXPECT.get(name, 0)
        if got.status == 0:
            assert got.counts == (0,) * 8
            assert np.abs(got.g.astype(np.float64) - f).max() <= eps


# 2-saddle walls (preserve_connectors): expected status per input of test_correction_bit_exact;
# without the finisher two of the inputs reach a fixed point with walls left
WALL_EXPECT = {"walls_geometry": (0, 2, 2), "walls_connectivity": (0, 2, 2), "walls_finish": (0, 0, 0)}
WALL_CASES = [(name, fl, WALL_EXPECT[name]) for name, fl in corpus.WALL_CASES]


@pytest.mark.parametrize("name,fl,expect", WALL_CASES, ids=[c[0] for c in WALL_CASES])
def test_correction_with_walls_bit_exact(oracle, dev, name, fl, expect):
    for (inp, f, g, eps, rng), status in zip(corpus.gm_inputs(), expect):
        print("correct", name, inp, f.shape, flush=True)
        got = _compare_run(oracle, dev, f, g, eps, rng, fl)
        assert got.status == status
        assert (got.trajectory.reshape(-1, 8)[:, 6] > 0).sum() > 1  # walls were found and repaired
        if status == 0:
            assert got.counts == (0,) * 8


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_wall_lists_bit_exact(oracle, dev, dtype):
    """Wall lists (face, first differing edge) and the fix_ds_wall edit set, on quantised g and
    on g whose critical cells already match f's."""
    seen = 0
    for name, f, g, eps, rng in corpus.wall_inputs(oracle):
        f = f.astype(dtype)
        g = g.astype(dtype)
        for edit_type in (0, 1):
            print("walls", name, dtype.__name__, edit_type, flush=True)
            ofl = oracle.flags(preserve_vpath=1, preserve_geometry=1, preserve_connectors=1, edit_type=edit_type)
            want = oracle.paths(f, g, ofl)
            got = dmtz_amd.walls(f, g, eps, rng, dmtz_amd.CorrectionFlags(edit_type=edit_type), ctx=dev,
                                 with_targets=True)
            assert sorted(zip(got["face"].tolist(), got["edge"].tolist())) == \
                sorted(zip(want["wall"].tolist(), want["wall_rec"].tolist()))
            want_t, n = oracle.wall_targets(f, g, eps, rng, ofl)
            assert n == len(got["face"])
            assert (got["targets"] == want_t).all()
            seen += n
    assert seen > 0


def test_correction_bit_exact_f64(oracle, dev):
    f = fields.gaussian_mixture((20, 18, 16), 6, seed=4, noise=0.02).astype(np.float64)
    eps, rng = fields.eps_from_rel(f, 2e-2)
    g = fields.quantize(f, eps)
    for fl in ({}, dict(preserve_vpath=1, preserve_geometry=1)):
        _compare_run(oracle, dev, f, g, eps, rng, fl)


def test_correction_c1_64cubed(oracle, dev):
    """The CPU-runnable workload: 64^3 Gaussian mixture (8 bumps), eps = 1e-3 range, geometry."""
    f = fields.gaussian_mixture((64, 64, 64), 8, seed=0)
    eps, rng = fields.eps_from_rel(f, 1e-3)
    g = fields.quantize(f, eps)
    got = _compare_run(oracle, dev, f, g, eps, rng, dict(preserve_vpath=1, preserve_geometry=1))
    assert got.counts == (0,) * 8


def test_nonconvergence_is_reported(dev):
    f = fields.gaussian_mixture((16, 16, 16), 6, seed=0, noise=0.02)
    eps, rng = fields.eps_from_rel(f, 2e-2)
    g = fields.quantize(f, eps)
    r = dmtz_amd.correct(f, g, eps, rng, dmtz_amd.CorrectionFlags(max_sweeps=1), ctx=dev)
    assert r.status == 1 and r.sweeps == 1 and any(c > 0 for c in r.counts)


def test_bad_arguments_raise(dev):
    f = np.zeros((4, 4, 2), np.float32)  # nx = 2 is below the minimum grid
    with pytest.raises(dmtz_amd.DmtzError):
        dmtz_amd.correct(f, f, 0.1, 1.0, ctx=dev)
    with pytest.raises(dmtz_amd.DmtzError):
        dmtz_amd.walls(f, f, ctx=dev)


@pytest.mark.parametrize("geometry", [1, 0])
def test_separatrix_lists_bit_exact(oracle, dev, geometry):
    """Separatrix detection lists and the path-repair edit set, on g whose critical cells
    already match f's (the state in which the reference runs these checks)."""
    import ctypes

    L = oracle.lib()
    P = ctypes.c_void_p
    I = ctypes.c_int64
    L.orc_path_targets_f32.argtypes = [P, P, I, I, I, ctypes.c_double, ctypes.c_double, P, P, P]
    for name, f, g, eps, rng in corpus.path_inputs(oracle):
        fl = dict(preserve_vpath=1, preserve_geometry=geometry)
        print("paths", name, geometry, flush=True)
        want = oracle.paths(f, g, oracle.flags(**fl))
        got = dmtz_amd.paths(f, g, dmtz_amd.CorrectionFlags(**fl), ctx=dev, with_targets=True)
        assert sorted(zip(got["ds"].tolist(), got["ds_rec"].tolist())) == \
            sorted(zip(want["ds"].tolist(), want["ds_rec"].tolist()))
        assert sorted(zip(got["as_"].tolist(), got["as_rec"].tolist())) == \
            sorted(zip(want["as_"].tolist(), want["as_rec"].tolist()))
        D, H, W = f.shape
        tg = np.zeros(W * H * D, np.uint8)
        c2 = np.zeros(2, np.int64)
        ofl = oracle.flags(**fl)
        L.orc_path_targets_f32(f.ctypes.data_as(P), g.ctypes.data_as(P), W, H, D, eps, rng,
                               ctypes.byref(ofl), tg.ctypes.data_as(P), c2.ctypes.data_as(P))
        assert (got["targets"].ravel() == tg.astype(bool)).all()


def test_classify_matches_oracle(oracle, dev):
    """dmtz_classify: critical-cell counts per type equal the -1 entries of the oracle's pairing
    arrays, the per-vertex masks list exactly the critical minima, and the critical edges the
    masks name are the oracle's."""
    for f in _gradient_cases():
        print("classify", f.shape, f.dtype, flush=True)
        pv, pe, pf, pt = oracle.gradient(f)
        counts, lo, hi = dmtz_amd.classify(f, ctx=dev, masks=True)
        assert counts == (int((pv == -1).sum()), int((pe == -1).sum()), int((pf == -1).sum()),
                          int((pt == -1).sum()))
        lo, hi = lo.reshape(-1), hi.reshape(-1)
        assert ((lo & 1).astype(bool) == (pv == -1)).all()
        bits = np.array([bin(int(x)).count("1") for x in lo]) + np.array([bin(int(x)).count("1") for x in hi])
        assert bits.sum() == sum(counts)
        D, H, W = f.shape
        dirs = [(1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1), (-1, 1, 0), (1, -1, 0),
                (0, 1, 1), (0, -1, -1), (-1, 0, 1), (1, 0, -1), (-1, 1, 1), (1, -1, -1)]
        edges = set()
        for v in np.nonzero((lo >> 1) & 0x3FFF)[0]:
            x, y, z = v % W, (v // W) % H, v // (W * H)
            for a in range(14):
                if (int(lo[v]) >> (1 + a)) & 1:
                    dx, dy, dz = dirs[a]
                    u = (x + dx) + (y + dy) * W + (z + dz) * W * H
                    edges.add(oracle.edge_id((W, H, D), int(v), int(u)))
        assert edges == set(np.nonzero(pe == -1)[0].tolist())


def test_batch_equals_single_corrections(dev):
    """dmtz_correct_batch over fields of different shapes gives each field exactly what a single
    correction gives."""
    items = []
    for seed, shape, rel in [(0, (16, 16, 16), 1e-2), (1, (24, 20, 12), 2e-2), (2, (13, 17, 11), 5e-2),
                             (3, (9, 30, 7), 3e-2)]:
        f = fields.gaussian_mixture(shape, 6, seed=seed, noise=0.02)
        eps, rng = fields.eps_from_rel(f, rel)
        items.append((f, fields.quantize(f, eps), eps, rng))
    fl = dmtz_amd.CorrectionFlags(preserve_vpath=1, preserve_geometry=1, finish=1)
    batch = dmtz_amd.correct_batch(items, fl, ctx=dev)
    for (f, g, eps, rng), b in zip(items, batch):
        one = dmtz_amd.correct(f, g, eps, rng, fl, ctx=dev)
        assert b.status == one.status and b.sweeps == one.sweeps and b.counts == one.counts
        assert b.g.tobytes() == one.g.tobytes()


DIRS14 = [(1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1), (-1, 1, 0), (1, -1, 0),
          (0, 1, 1), (0, -1, -1), (-1, 0, 1), (1, 0, -1), (-1, 1, 1), (1, -1, -1)]


def _local_cells():
    """Star triangles and tets of the local-cell numbering (slot pairs / triples, adjacent slots),
    in the order of the classifier's masks."""
    dirs = set(DIRS14)

    def adj(i, j):
        return tuple(a - b for a, b in zip(DIRS14[i], DIRS14[j])) in dirs

    tris = [(i, j) for i in range(14) for j in range(i + 1, 14) if adj(i, j)]
    tets = [(i, j, k) for i in range(14) for j in range(i + 1, 14) for k in range(j + 1, 14)
            if adj(i, j) and adj(i, k) and adj(j, k)]
    return tris, tets


def _critical_cells(lo, hi, shape):
    """Critical cells named by per-owner masks, as sorted rows of vertex ids per type."""
    D, H, W = shape
    lo, hi = lo.reshape(-1), hi.reshape(-1)
    off = [dx + dy * W + dz * W * H for dx, dy, dz in DIRS14]
    tris, tets = _local_cells()
    assert len(tris) == 36 and len(tets) == 24

    def rows(groups):
        if not groups:
            return np.zeros((0, 1), np.int64)
        r = np.sort(np.concatenate(groups, axis=0), axis=1)
        return r[np.lexsort(r.T[::-1])]

    out = [np.sort(np.nonzero(lo & 1)[0])]
    out.append(rows([np.stack([v, v + off[a]], 1) for a in range(14)
                     for v in [np.nonzero((lo >> np.uint64(1 + a)) & np.uint64(1))[0]]]))
    out.append(rows([np.stack([v, v + off[i], v + off[j]], 1) for t, (i, j) in enumerate(tris)
                     for v in [np.nonzero((lo >> np.uint64(15 + t)) & np.uint64(1))[0]]]))
    out.append(rows([np.stack([v, v + off[i], v + off[j], v + off[k]], 1) for q, (i, j, k) in enumerate(tets)
                     for v in [np.nonzero((hi >> np.uint32(q)) & np.uint32(1))[0]]]))
    return out


@pytest.mark.timeout(400)
def test_fbm_512_converges_with_f_critical_cells(dev):
    """The 512^3 fBm workload (P(k) ~ k^-3 on [0, 1], eps = 1e-3 range, separatrix geometry,
    finisher), too large for the oracle: the run ends at zero mismatches, every value stays within
    eps, and the classifier finds exactly f's critical cells in the corrected field (compared as
    cells: a cell's owner, its highest vertex, may differ between f and g)."""
    f = fields.fbm((512, 512, 512), seed=0)
    eps, rng = fields.eps_from_rel(f, 1e-3)
    g = fields.quantize(f, eps)
    fl = dmtz_amd.CorrectionFlags(preserve_vpath=1, preserve_geometry=1, finish=1, max_sweeps=2000)
    print("fbm 512^3 correction", flush=True)
    r = dmtz_amd.correct(f, g, eps, rng, fl, ctx=dev)
    assert r.status == 0 and r.counts == (0,) * 8
    assert np.abs(r.g.astype(np.float64) - f).max() <= eps
    cf, lof, hif = dmtz_amd.classify(f, ctx=dev, masks=True)
    cg, log_, hig = dmtz_amd.classify(r.g, ctx=dev, masks=True)
    assert cf == cg
    for a, b in zip(_critical_cells(lof, hif, f.shape), _critical_cells(log_, hig, f.shape)):
        assert np.array_equal(a, b)


def test_critical_cell_sets_small(oracle, dev):
    """The cell-set form of the classifier's masks (used at 512^3) against the oracle's critical
    sets on small fields, so the comparison above is itself checked."""
    for f in _gradient_cases()[:6]:
        D, H, W = f.shape
        pv, pe, pf, pt = oracle.gradient(f)
        _, lo, hi = dmtz_amd.classify(f, ctx=dev, masks=True)
        cells = _critical_cells(lo, hi, f.shape)
        dims = (W, H, D)
        assert cells[0].tolist() == np.nonzero(pv == -1)[0].tolist()
        want_e = sorted(tuple(sorted(oracle.edge_verts(dims, int(e)))) for e in np.nonzero(pe == -1)[0])
        assert [tuple(r) for r in cells[1].tolist()] == want_e
        want_t = sorted(tuple(sorted(oracle.tri_verts(dims, int(t)))) for t in np.nonzero(pf == -1)[0])
        assert [tuple(r) for r in cells[2].tolist()] == want_t
        want_q = sorted(tuple(sorted(oracle.tet_verts(dims, int(q)))) for q in np.nonzero(pt == -1)[0])
        assert [tuple(r) for r in cells[3].tolist()] == want_q


def test_batch_error_leaves_unrun_fields_marked(dev):
    """A field that fails (nx = 2 is below the minimum grid) in the middle of a batch: its status
    is the argument error, fields that ran report their own results, and fields that never
    started report NOT_RUN with their g untouched."""
    items = []
    for seed, shape in [(0, (12, 11, 10)), (1, (9, 8, 2)), (2, (10, 9, 8)), (3, (11, 10, 9))]:
        f = fields.gaussian_mixture(shape, 4, seed=seed, noise=0.02)
        eps, rng = fields.eps_from_rel(f, 2e-2)
        items.append((f, fields.quantize(f, eps), eps, rng))
    fl = dmtz_amd.CorrectionFlags()
    with pytest.raises(dmtz_amd.DmtzError) as ei:
        dmtz_amd.correct_batch(items, fl, ctx=dev)
    res = ei.value.results
    assert res[1].status == -1  # DMTZ_E_ARG
    for (f, g, eps, rng), r in zip(items, res):
        assert r.status in (0, -1, dmtz_amd.NOT_RUN)
        if r.status == dmtz_amd.NOT_RUN:
            assert r.g.tobytes() == g.tobytes() and r.sweeps == 0
        elif r.status == 0:
            one = dmtz_amd.correct(f, g, eps, rng, fl, ctx=dev)
            assert r.g.tobytes() == one.g.tobytes() and r.sweeps == one.sweeps
