"""The product's device-agnostic headers (lower_star.h, gradient.h, mismatch.h, repair.h), built
for the host with g++ as a test harness, against the oracle: pairing arrays bit-exact, mismatch
counts and repair edit sets identical. This runs the same classifier and repair code the HIP
kernels run, without a GPU."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

from dmtz_amd import fields

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SRC = os.path.join(HERE, "native", "hostcheck.cpp")
SO = os.path.join(HERE, "native", "libhostcheck.so")
HDRS = [os.path.join(ROOT, "dmtz_amd", "csrc", h) for h in
        ("stencil.h", "lower_star.h", "gradient.h", "mismatch.h", "repair.h", "edit.h")]

P = ctypes.c_void_p
I = ctypes.c_int64


@pytest.fixture(scope="module")
def hc():
    so = os.environ.get("DMTZ_HOSTCHECK_SO")  # a sanitizer build (test_hostcheck_sanitized.py)
    if not so:
        so = SO
        newest = max(os.path.getmtime(p) for p in HDRS + [SRC])
        if not os.path.exists(SO) or os.path.getmtime(SO) < newest:
            subprocess.check_call(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off",
                                   "-I" + os.path.join(ROOT, "dmtz_amd", "csrc"), "-o", SO, SRC])
    L = ctypes.CDLL(so)
    L.hc_gradient_f32.argtypes = [P, I, I, I, P, P, P, P]
    L.hc_gradient_f64.argtypes = [P, I, I, I, P, P, P, P]
    for t in ("f32", "f64"):
        getattr(L, "hc_targets_" + t).argtypes = [P, P, I, I, I, ctypes.c_double, ctypes.c_double,
                                                 ctypes.c_int, ctypes.c_int, P, P]
    return L


def _p(a):
    return a.ctypes.data_as(P)


def hc_gradient(L, f, oracle):
    D, H, W = f.shape
    outs = [np.empty(n, np.int64) for n in oracle.counts(W, H, D)]
    fn = L.hc_gradient_f32 if f.dtype == np.float32 else L.hc_gradient_f64
    fn(_p(f), W, H, D, *[_p(o) for o in outs])
    return outs


def gradient_cases():
    cases = []
    for seed in range(4):
        for shape in [(7, 6, 5), (5, 8, 9), (9, 9, 9), (4, 3, 6)]:
            cases.append(fields.gaussian_mixture(shape, 3, seed=seed, noise=0.05))
    rng = np.random.default_rng(1)
    for k in range(4):
        cases.append(rng.integers(0, 3, size=(6, 7, 8)).astype(np.float32))
        cases.append(rng.integers(0, 2, size=(5, 5, 5)).astype(np.float64))
    cases.append(np.zeros((3, 4, 5), np.float32))
    return cases


def test_lower_star_gradient_matches_oracle(hc, oracle):
    for f in gradient_cases():
        A = oracle.gradient(f)
        B = hc_gradient(hc, f, oracle)
        for name, a, b in zip(("vertices", "edges", "faces", "tets"), A, B):
            assert (a == b).all(), (f.shape, f.dtype, name, np.nonzero(a != b)[0][:5])


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("edit_type", [0, 1])
def test_repair_edit_set_matches_oracle(hc, oracle, dtype, edit_type):
    for seed in range(5):
        for shape, rel in [((8, 9, 10), 1e-2), ((12, 12, 12), 3e-2), ((6, 7, 9), 5e-2)]:
            f = fields.gaussian_mixture(shape, 4, seed=seed, noise=0.02).astype(dtype)
            eps, rng = fields.eps_from_rel(f, rel)
            g = fields.quantize(f, eps)
            want, wc = oracle.targets(f, g, eps, rng, oracle.flags(edit_type=edit_type))
            D, H, W = shape
            tg = np.zeros(W * H * D, np.uint8)
            c4 = np.zeros(4, np.int64)
            fn = hc.hc_targets_f32 if dtype == np.float32 else hc.hc_targets_f64
            fn(_p(f), _p(g), W, H, D, eps, rng, 1, edit_type, _p(tg), _p(c4))
            assert tuple(c4) == wc
            assert (tg.reshape(shape).astype(bool) == want).all()


def test_precomputed_ascending_paths_equal_the_walk(hc, oracle):
    """The ascending geometry check over steps precomputed on f alone (Repair::as_fpath, used by
    the device's incremental path checks) finds exactly what the direct walk (Repair::as_walk,
    is_false of main_3d.cu:4437-4497) finds, for every branch."""
    hc.hc_as_two_ways_f32.argtypes = [P, P, I, I, I, P, P]
    hc.hc_as_two_ways_f32.restype = I
    found = 0
    for seed in range(4):
        for shape, rel in [((12, 12, 12), 1e-2), ((10, 14, 9), 3e-2), ((16, 16, 16), 5e-3)]:
            f = fields.gaussian_mixture(shape, 5, seed=seed, noise=0.02)
            eps, rng = fields.eps_from_rel(f, rel)
            g0 = fields.quantize(f, eps)
            g1 = oracle.correct(f, g0, eps, rng, oracle.flags())["g"]  # critical cells only
            for g in (g0, g1):
                D, H, W = shape
                cap = 2 * 36 * W * H * D
                ow, os_ = np.zeros(cap, np.int64), np.zeros(cap, np.int64)
                nb = hc.hc_as_two_ways_f32(_p(f), _p(g), W, H, D, _p(ow), _p(os_))
                assert nb > 0
                assert (ow[: 2 * nb] == os_[: 2 * nb]).all()
                found += int((ow[: 2 * nb: 2] >= 0).sum())
    assert found > 0


def test_fast_division_matches_native(hc):
    """The multiply-high division the ID decoders use (stencil.h FastDiv / divmod) equals native
    64-bit division and C remainder semantics, negative numerators included."""
    hc.hc_divmod_check.argtypes = [I]
    hc.hc_divmod_check.restype = I
    assert hc.hc_divmod_check(200) == 0


@pytest.mark.parametrize("geometry", [1, 0])
def test_separatrix_lists_match_oracle(hc, oracle, geometry):
    """Separatrix detection lists (geometry and connectivity mode) from the host build of the
    device headers against the oracle, on g whose critical cells already match f's."""
    hc.hc_paths_f32.argtypes = [P, P, I, I, I, ctypes.c_int, P, P, P, P, P]
    for seed in range(3):
        for shape, rel in [((12, 12, 12), 1e-2), ((10, 14, 9), 3e-2), ((13, 17, 11), 5e-2)]:
            f = fields.gaussian_mixture(shape, 5, seed=seed, noise=0.02)
            eps, rng = fields.eps_from_rel(f, rel)
            g = oracle.correct(f, fields.quantize(f, eps), eps, rng, oracle.flags())["g"]
            want = oracle.paths(f, g, oracle.flags(preserve_vpath=1, preserve_geometry=geometry))
            D, H, W = shape
            N, E, F, T = oracle.counts(W, H, D)
            c2 = np.zeros(2, np.int64)
            ds_id, ds_rec = np.zeros(E, np.int64), np.zeros(E, np.int64)
            as_id, as_rec = np.zeros(2 * F, np.int64), np.zeros(2 * F, np.int64)
            rc = hc.hc_paths_f32(_p(f), _p(g), W, H, D, geometry, _p(c2), _p(ds_id), _p(ds_rec), _p(as_id),
                                 _p(as_rec))
            assert rc == 0
            nds, nas = int(c2[0]), int(c2[1])
            assert sorted(zip(ds_id[:nds].tolist(), ds_rec[:nds].tolist())) == \
                sorted(zip(want["ds"].tolist(), want["ds_rec"].tolist()))
            assert sorted(zip(as_id[:nas].tolist(), as_rec[:nas].tolist())) == \
                sorted(zip(want["as_"].tolist(), want["as_rec"].tolist()))


def test_decoders_match_oracle_on_every_id(hc, oracle):
    """Triangle and tet decoders on in-range, negative and past-the-end ids: the device decoders
    (stencil.h) decode exactly as the oracle's restatement of the reference, whose arithmetic
    gives a negative id a definite pattern."""
    hc.hc_decode.argtypes = [I, I, I, P, I, P, I, P]
    for dims in [(5, 6, 7), (3, 2, 2), (11, 17, 13)]:
        W, H, D = dims
        N, E, F, T = oracle.counts(W, H, D)
        tris = np.array(list(range(-13, 40)) + list(range(F - 20, F + 20)), np.int64)
        tets = np.array(list(range(-13, 40)) + list(range(T - 20, T + 20)), np.int64)
        out = np.zeros(3 * len(tris) + 4 * len(tets), np.int64)
        hc.hc_decode(W, H, D, _p(tris), len(tris), _p(tets), len(tets), _p(out))
        got_t = out[: 3 * len(tris)].reshape(-1, 3)
        got_q = out[3 * len(tris):].reshape(-1, 4)
        for k, t in enumerate(tris):
            assert tuple(got_t[k]) == oracle.tri_verts(dims, int(t)), (dims, t)
        for k, q in enumerate(tets):
            assert tuple(got_q[k]) == oracle.tet_verts(dims, int(q)), (dims, q)


def test_critical_masks_scattered_equal_gathered(hc):
    """The f-critical masks of all local cells, spread from each cell's owner and gathered from
    the neighbours' owned critical cells (the device's k_critf_gather), equal the masks found by
    pairing lookups at the owners (local_crit)."""
    for name in ("hc_critf_two_ways_f32", "hc_critf_two_ways_f64"):
        getattr(hc, name).argtypes = [P, I, I, I]
        getattr(hc, name).restype = I
    cases = gradient_cases() + [fields.gaussian_mixture((17, 19, 23), 6, seed=9, noise=0.02)]
    for f in cases:
        D, H, W = f.shape
        fn = hc.hc_critf_two_ways_f32 if f.dtype == np.float32 else hc.hc_critf_two_ways_f64
        assert fn(_p(f), W, H, D) == 0, (f.shape, f.dtype)


def test_walls_match_oracle(hc, oracle):
    """Wall lists and the fix_ds_wall edit set from the host build of Repair::wall_walk (the
    code the wall kernel runs) against the oracle, on quantised g and on g whose critical cells
    already match f's."""
    hc.hc_walls_f32.argtypes = [P, P, I, I, I, ctypes.c_double, ctypes.c_double, ctypes.c_int, P, P, P]
    hc.hc_walls_f32.restype = I
    seen = 0
    for seed, shape, rel in [(2, (16, 16, 16), 5e-3), (3, (13, 17, 11), 5e-2)]:
        f = fields.gaussian_mixture(shape, 6, seed=seed, noise=0.02)
        eps, rng = fields.eps_from_rel(f, rel)
        g0 = fields.quantize(f, eps)
        g1 = oracle.correct(f, g0, eps, rng, oracle.flags())["g"]
        D, H, W = shape
        F = oracle.counts(W, H, D)[2]
        for g in (g0, g1):
            for edit_type in (0, 1):
                ofl = oracle.flags(preserve_vpath=1, preserve_geometry=1, preserve_connectors=1, edit_type=edit_type)
                want = oracle.paths(f, g, ofl)
                face, edge = np.zeros(F, np.int64), np.zeros(F, np.int64)
                tg = np.zeros(f.size, np.uint8)
                n = hc.hc_walls_f32(_p(f), _p(g), W, H, D, eps, rng, edit_type, _p(face), _p(edge), _p(tg))
                assert n >= 0
                assert sorted(zip(face[:n].tolist(), edge[:n].tolist())) == \
                    sorted(zip(want["wall"].tolist(), want["wall_rec"].tolist()))
                want_t, nw = oracle.wall_targets(f, g, eps, rng, ofl)
                assert nw == n
                assert (tg.astype(bool).reshape(f.shape) == want_t).all()
                seen += n
    assert seen > 0
