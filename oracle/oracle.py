"""ctypes wrapper of the C++ oracle (liboracle.so) — TEST INFRASTRUCTURE ONLY.

Only the test suite, ``__graft_entry__.smoke()`` and the CPU-baseline leg of ``bench.py`` may
import this module, and only as the checker; the shipped HIP path never touches it.

The oracle restates the reference's topology-correction path of main_3d.cu (dense per-simplex
pairing arrays keyed by the reference's global simplex IDs); see dmtz_oracle.cpp for the
line-by-line citations.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

I64 = ctypes.c_int64
P = ctypes.c_void_p


class OrcFlags(ctypes.Structure):
    # Field order of the reference's argv flags, main_3d.cu:6264-6271, plus a sweep cap and the
    # finisher switch (restore f around a stalled edit set; no counterpart in the reference).
    _fields_ = [
        ("preserve_saddles", ctypes.c_int32),
        ("preserve_vpath", ctypes.c_int32),
        ("preserve_geometry", ctypes.c_int32),
        ("filtration", ctypes.c_int32),
        ("preserve_connectors", ctypes.c_int32),
        ("preserve_type", ctypes.c_int32),
        ("edit_type", ctypes.c_int32),
        ("max_sweeps", ctypes.c_int32),
        ("finish", ctypes.c_int32),
    ]


def build(force=False):
    so = os.path.join(_HERE, "liboracle.so")
    src = os.path.join(_HERE, "dmtz_oracle.cpp")
    if force or not os.path.exists(so) or os.path.getmtime(so) < os.path.getmtime(src):
        subprocess.check_call(["make", "-s", "-C", _HERE, "-B" if force else "all"])
    return so


def lib():
    global _LIB
    if _LIB is None:
        so = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(so):
            build()
        L = ctypes.CDLL(so)
        L.orc_edge_id.restype = I64
        L.orc_tri_id.restype = I64
        L.orc_tet_id.restype = I64
        for name in ("orc_edge_id",):
            getattr(L, name).argtypes = [I64] * 5
        L.orc_tri_id.argtypes = [I64] * 6
        L.orc_tet_id.argtypes = [I64] * 7
        L.orc_edge_verts.argtypes = [I64, I64, I64, I64, P]
        L.orc_tri_verts.argtypes = [I64, I64, I64, I64, P]
        L.orc_tet_verts.argtypes = [I64, I64, I64, I64, P]
        L.orc_cell_tetra.argtypes = [I64, I64, I64, I64, ctypes.c_int, P]
        L.orc_counts.argtypes = [I64, I64, I64, P]
        for t in ("f32", "f64"):
            getattr(L, "orc_gradient_" + t).argtypes = [P, I64, I64, I64, ctypes.c_int, P, P, P, P]
            getattr(L, "orc_targets_" + t).argtypes = [P, P, I64, I64, I64, ctypes.c_double,
                                                      ctypes.c_double, P, P, P]
            getattr(L, "orc_paths_" + t).argtypes = [P, P, I64, I64, I64, P, P, P, P, P, P, P, P]
            getattr(L, "orc_wall_targets_" + t).argtypes = [P, P, I64, I64, I64, ctypes.c_double,
                                                           ctypes.c_double, P, P, P]
            getattr(L, "orc_correct_" + t).argtypes = [P, P, I64, I64, I64, ctypes.c_double,
                                                      ctypes.c_double, P, P, P, P, P, I64, P, P]
        L.orc_set_threads.argtypes = [ctypes.c_int]
        L.orc_fix_one_f32.argtypes = [P, P, I64, I64, I64, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                      ctypes.c_int, I64, P]
        L.orc_coverage.argtypes = [ctypes.c_int, P, ctypes.c_int]
        L.orc_coverage_name.argtypes = [ctypes.c_int]
        L.orc_coverage_name.restype = ctypes.c_char_p
        L.orc_encode3.argtypes = [P, I64, P, I64]
        L.orc_encode3.restype = I64
        L.orc_rle_bitmap.argtypes = [P, I64, P, I64]
        L.orc_rle_bitmap.restype = I64
        L.orc_cost_streams.argtypes = [P, P, P, P, I64, ctypes.c_double, P, P, I64, P]
        L.orc_cost_streams.restype = I64
        L.orc_original_cost_streams.argtypes = [P, P, I64, P, P, I64]
        L.orc_original_cost_streams.restype = I64
        _LIB = L
    return _LIB


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


def _tag(a):
    if a.dtype == np.float32:
        return "f32"
    if a.dtype == np.float64:
        return "f64"
    raise TypeError("field must be float32 or float64")


def coverage(enable=None):
    """Branch-hit counters of the restated repair rules and separatrix checks: enable=True clears
    and starts them, False stops them; returns {branch name and cited lines: hits}."""
    L = lib()
    n = L.orc_coverage(-1, None, 0)
    if enable is not None:
        L.orc_coverage(1 if enable else 0, None, 0)
    out = np.zeros(n, np.int64)
    L.orc_coverage(-1, _p(out), n)
    return {L.orc_coverage_name(k).decode(): int(out[k]) for k in range(n)}


def set_threads(n):
    return lib().orc_set_threads(int(n))


def counts(W, H, D):
    out = np.zeros(4, np.int64)
    lib().orc_counts(W, H, D, _p(out))
    return tuple(int(x) for x in out)


def edge_id(dims, a, b):
    return lib().orc_edge_id(*dims, a, b)


def edge_verts(dims, e):
    out = np.zeros(2, np.int64)
    lib().orc_edge_verts(*dims, e, _p(out))
    return tuple(int(x) for x in out)


def tri_id(dims, a, b, c):
    return lib().orc_tri_id(*dims, a, b, c)


def tri_verts(dims, t):
    out = np.zeros(3, np.int64)
    lib().orc_tri_verts(*dims, t, _p(out))
    return tuple(int(x) for x in out)


def tet_id(dims, a, b, c, d):
    return lib().orc_tet_id(*dims, a, b, c, d)


def tet_verts(dims, q):
    out = np.zeros(4, np.int64)
    lib().orc_tet_verts(*dims, q, _p(out))
    return tuple(int(x) for x in out)


def cell_tetra(dims, t, literal=True):
    out = np.zeros(2, np.int64)
    lib().orc_cell_tetra(*dims, t, 1 if literal else 0, _p(out))
    return tuple(int(x) for x in out)


def flags(preserve_saddles=1, preserve_vpath=0, preserve_geometry=0, filtration=0,
          preserve_connectors=0, preserve_type=0, edit_type=0, max_sweeps=100000, finish=0):
    return OrcFlags(preserve_saddles, preserve_vpath, preserve_geometry, filtration,
                    preserve_connectors, preserve_type, edit_type, max_sweeps, finish)


def gradient(field, literal=False):
    """Pairing arrays (pv, pe, pf, pt) of the reference's five-pass gradient for one field.

    ``field`` is indexed [z, y, x] (x fastest in memory, as main_3d.cu:1945-1947)."""
    field = np.ascontiguousarray(field)
    D, H, W = field.shape
    N, E, F, T = counts(W, H, D)
    pv = np.empty(N, np.int64)
    pe = np.empty(E, np.int64)
    pf = np.empty(F, np.int64)
    pt = np.empty(T, np.int64)
    getattr(lib(), "orc_gradient_" + _tag(field))(_p(field), W, H, D, 1 if literal else 0,
                                                  _p(pv), _p(pe), _p(pf), _p(pt))
    return pv, pe, pf, pt


def targets(f, g, eps, rng, fl=None):
    """One repair phase on (f, g): the set of vertices the fix rules edit, and the 4 counts."""
    f = np.ascontiguousarray(f)
    g = np.ascontiguousarray(g, dtype=f.dtype)
    D, H, W = f.shape
    fl = fl or flags()
    tg = np.zeros(W * H * D, np.uint8)
    c4 = np.zeros(4, np.int64)
    getattr(lib(), "orc_targets_" + _tag(f))(_p(f), _p(g), W, H, D, float(eps), float(rng),
                                             ctypes.byref(fl), _p(tg), _p(c4))
    return tg.astype(bool).reshape(f.shape), tuple(int(x) for x in c4)


def paths(f, g, fl):
    """Separatrix mismatch lists (ds, as, ds_wall) on (f, g) as they stand."""
    f = np.ascontiguousarray(f)
    g = np.ascontiguousarray(g, dtype=f.dtype)
    D, H, W = f.shape
    N, E, F, T = counts(W, H, D)
    c3 = np.zeros(3, np.int64)
    ds_l = np.zeros(E, np.int64)
    ds_r = np.zeros(E, np.int64)
    as_l = np.zeros(F * 2, np.int64)
    as_r = np.zeros(F * 2, np.int64)
    w_l = np.zeros(F, np.int64)
    w_r = np.zeros(F, np.int64)
    rc = getattr(lib(), "orc_paths_" + _tag(f))(_p(f), _p(g), W, H, D, ctypes.byref(fl), _p(c3),
                                                 _p(ds_l), _p(ds_r), _p(as_l), _p(as_r), _p(w_l),
                                                 _p(w_r))
    n_ds, n_as, n_w = (int(x) for x in c3)
    return dict(rc=rc, ds=ds_l[:n_ds], ds_rec=ds_r[:n_ds], as_=as_l[:n_as], as_rec=as_r[:n_as],
                wall=w_l[:n_w], wall_rec=w_r[:n_w])


def fix_one(f, g, eps, rng, cell_type, cell, edit_type=0):
    """The repair rule of one mismatched cell (cell_type 0..3: minimum, 1-saddle, 2-saddle,
    maximum; cell: its reference simplex ID) on float32 (f, g): the sorted vertices of its edit
    set, or None when the cell is not mismatched."""
    f = np.ascontiguousarray(f, dtype=np.float32)
    g = np.ascontiguousarray(g, dtype=np.float32)
    D, H, W = f.shape
    tg = np.zeros(W * H * D, np.uint8)
    rc = lib().orc_fix_one_f32(_p(f), _p(g), W, H, D, float(eps), float(rng), int(edit_type), int(cell_type),
                               int(cell), _p(tg))
    return sorted(np.nonzero(tg)[0].tolist()) if rc == 1 else None


def wall_targets(f, g, eps, rng, fl=None):
    """Wall list size and the edit set of fix_ds_wall over it, on (f, g) as they stand."""
    f = np.ascontiguousarray(f)
    g = np.ascontiguousarray(g, dtype=f.dtype)
    D, H, W = f.shape
    fl = fl or flags()
    tg = np.zeros(W * H * D, np.uint8)
    n = np.zeros(1, np.int64)
    getattr(lib(), "orc_wall_targets_" + _tag(f))(_p(f), _p(g), W, H, D, float(eps), float(rng),
                                                  ctypes.byref(fl), _p(tg), _p(n))
    return tg.astype(bool).reshape(f.shape), int(n[0])


def correct(f, g, eps, rng, fl=None, traj_cap=4096):
    """Run the reference correction loop.

    Returns dict(g, counts, counter, sweeps, status, traj, phases, stalls, finished, radius); status
    0 converged, 1 max_sweeps reached, 2 stalled (fixed point with mismatches left), -2 capped walk."""
    f = np.ascontiguousarray(f)
    g = np.array(g, dtype=f.dtype, copy=True, order="C")
    D, H, W = f.shape
    fl = fl or flags()
    c8 = np.zeros(8, np.int64)
    counter = np.zeros(W * H * D, np.int32)
    sweeps = np.zeros(1, np.int64)
    traj = np.zeros(traj_cap * 8, np.int64)
    phases = np.zeros(8, np.float64)
    info = np.zeros(4, np.int64)
    rc = getattr(lib(), "orc_correct_" + _tag(f))(_p(f), _p(g), W, H, D, float(eps), float(rng),
                                                   ctypes.byref(fl), _p(c8), _p(counter),
                                                   _p(sweeps), _p(traj), traj_cap, _p(phases), _p(info))
    ns = int(sweeps[0])
    return dict(g=g, counts=tuple(int(x) for x in c8), counter=counter.reshape(f.shape),
                sweeps=ns, status=rc, traj=traj[:min(ns, traj_cap) * 8].reshape(-1, 8),
                phases=dict(zip(("gradient_pairing", "get_f_cp", "get_f_path", "fix_f_cp",
                                 "fix_f_vpath"), phases[:5].tolist())),
                stalls=int(info[0]), finished=int(info[1]), radius=int(info[2]))


def encode3(values):
    """encodeTo3BitBitmap (main_3d.cu:1446-1476) of int values 0..6; None when one is outside."""
    v = np.ascontiguousarray(values, dtype=np.int32)
    out = np.zeros((3 * v.size + 7) // 8 + 1, np.uint8)
    k = lib().orc_encode3(_p(v), v.size, _p(out), out.size)
    return None if k < 0 else out[:k].tobytes()


def rle_bitmap(bitmap):
    """runLengthEncodeBitmap (main_3d.cu:1399-1426): (bit, run length) pairs, int32, shape (runs, 2)."""
    b = np.frombuffer(bytes(bitmap), np.uint8).copy()
    L = lib()
    runs = L.orc_rle_bitmap(_p(b), b.size, None, 0)
    out = np.zeros((max(runs, 1), 2), np.int32)
    L.orc_rle_bitmap(_p(b), b.size, _p(out), runs)
    return out[:runs]


def cost_streams(f, g0, g1, counter, bound):
    """The index and value streams of cost (main_3d.cu:1494-1510, float64 fields): the per-vertex
    indices (counter of changed vertices), their bitmap, and the index-6 values; plus the changed
    vertex count."""
    f, g0, g1 = (np.ascontiguousarray(a, np.float64).ravel() for a in (f, g0, g1))
    c = np.ascontiguousarray(counter, np.uint8).ravel()
    idx = np.zeros(f.size, np.int32)
    edits = np.zeros(f.size, np.float64)
    m = ctypes.c_int64(0)
    cnt = lib().orc_cost_streams(_p(f), _p(g0), _p(g1), _p(c), f.size, float(bound), _p(idx), _p(edits), edits.size,
                                 ctypes.byref(m))
    return dict(index=idx, bitmap=encode3(idx), edits=edits[: m.value].copy(), edited=int(cnt))


def original_cost_streams(g0, g1):
    """The index differences and values of original_cost (main_3d.cu:1649-1671)."""
    g0, g1 = (np.ascontiguousarray(a, np.float64).ravel() for a in (g0, g1))
    diffs = np.zeros(g0.size, np.int32)
    edits = np.zeros(g0.size, np.float64)
    m = lib().orc_original_cost_streams(_p(g0), _p(g1), g0.size, _p(diffs), _p(edits), g0.size)
    return dict(diffs=diffs[:m].copy(), edits=edits[:m].copy(), edited=int(m))
