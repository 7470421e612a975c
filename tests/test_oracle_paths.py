"""Pins of the oracle's separatrix checks by an independent serial restatement.

Each check below is restated here in plain Python from the reference's own text, over the
pairing arrays of f and g (the oracle's gradients, themselves pinned by the literal 14^k search,
the Euler characteristic and matching validity in test_oracle.py) and the simplex-ID maps (pinned
by the bijection tests there). Face cofacets are enumerated here from the 14-neighbour stencil in
the reference's loop order, not through the oracle's cofacet routine. The oracle's lists must
equal these on tiny grids, for quantised g and for g whose critical cells already match f's (the
state in which the reference runs these checks).

  get_false_ds_vpath   main_3d.cu:4837-4959  (descending geometry)
  is_false             main_3d.cu:4437-4497  (ascending walk, first divergent tet)
  get_false_as_vpath   main_3d.cu:4540-4610  (ascending geometry, cofacet order)
  get_paired_id        main_3d.cu:4499-4537  (maximum reached from a tet)
  get_orginal_saddle_min_pairs / _max_pairs   main_3d.cu:4613-4759  (f labels)
  get_false_ds_vpath_connectivity  main_3d.cu:4961-5062
  get_false_as_vpath_connectivity  main_3d.cu:4761-4835 (labels indexed by face id: the
                                   reference's inner loop shadows the face index there)
"""
import numpy as np
import pytest

from dmtz_amd import fields

# main_3d.cu:44-51
DIRS = [(1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1), (-1, 1, 0), (1, -1, 0),
        (0, 1, 1), (0, -1, -1), (-1, 0, 1), (1, 0, -1), (-1, 1, 1), (1, -1, -1)]


class Walker:
    def __init__(self, oracle, f, g):
        D, H, W = f.shape
        self.O = oracle
        self.W, self.H, self.D = W, H, D
        self.dims = (W, H, D)
        self.pv = [None, None]
        self.pe = [None, None]
        self.pf = [None, None]
        self.pt = [None, None]
        for k, fld in enumerate((f, g)):
            self.pv[k], self.pe[k], self.pf[k], self.pt[k] = oracle.gradient(fld)
        self._cof = {}

    def xyz(self, v):
        return v % self.W, (v // self.W) % self.H, v // (self.W * self.H)

    def inb(self, x, y, z):
        return 0 <= x < self.W and 0 <= y < self.H and 0 <= z < self.D

    def adjacent(self, a, b):
        ax, ay, az = self.xyz(a)
        bx, by, bz = self.xyz(b)
        return (bx - ax, by - ay, bz - az) in DIRS

    def face_cofacets(self, t):
        """Tets of face t in the order of the direction loop over its highest-index vertex
        (main_3d.cu:4567-4606): v4 = v1 + DIRS[i] that is a neighbour of v2 and v3."""
        if t in self._cof:
            return self._cof[t]
        v1, v2, v3 = self.O.tri_verts(self.dims, t)
        x, y, z = self.xyz(v1)
        out = []
        for dx, dy, dz in DIRS:
            if not self.inb(x + dx, y + dy, z + dz):
                continue
            n = (x + dx) + (y + dy) * self.W + (z + dz) * self.W * self.H
            if n in (v1, v2, v3) or not self.adjacent(v2, n) or not self.adjacent(v3, n):
                continue
            out.append(self.O.tet_id(self.dims, v1, v2, v3, n))
        self._cof[t] = out
        return out

    @staticmethod
    def at(a, i):
        return int(a[i]) if 0 <= i < len(a) else -1

    # is_false main_3d.cu:4437-4497 -> the divergent tet or None
    def is_false(self, tet):
        cur = tet
        for _ in range(10 ** 6):
            old = cur
            t = self.at(self.pt[0], cur)
            dt = self.at(self.pt[1], cur)
            if t == -1:
                return None
            if t != dt:
                return cur
            p = self.at(self.pf[0], t)
            if p == -1 or self.at(self.pt[0], p) != t:
                return None
            cof = self.face_cofacets(t)
            nxt = [q for q in cof if q != cur]
            cur = nxt[0] if nxt else -1
            if cur == old or cur == -1:
                return None
        raise AssertionError("ascending walk did not end")

    # get_paired_id main_3d.cu:4499-4537 (k = 0: f arrays, 1: g arrays)
    def paired_max(self, tet, k):
        cur = tet
        for _ in range(10 ** 6):
            old = cur
            t = self.at(self.pt[k], cur)
            if t == -1:
                return cur
            p = self.at(self.pf[k], t)
            if p == -1 or self.at(self.pt[k], p) != t:
                return cur
            nxt = [q for q in self.face_cofacets(t) if q != cur]
            cur = nxt[0] if nxt else -1
            if cur == old or cur == -1:
                return cur
        raise AssertionError("label walk did not end")

    # the do-while descent of get_orginal_saddle_min_pairs main_3d.cu:4631-4684 over array k
    def min_label(self, v, k):
        for _ in range(10 ** 6):
            e = self.at(self.pv[k], v)
            if e == -1:
                return v
            a, b = self.O.edge_verts(self.dims, e)
            v = b if a == v else a
        raise AssertionError("descending walk did not end")

    def f_critical_edges(self):
        return [int(i) for i in np.nonzero(self.pe[0] == -1)[0]]

    def f_critical_faces(self):
        return [int(i) for i in np.nonzero(self.pf[0] == -1)[0]]

    # get_false_ds_vpath main_3d.cu:4837-4959: dec_current_id never advances (:4849), so the
    # second pass of each do-while leaves through v != dec_v (:4861) and only the start vertex of
    # each endpoint walk is compared.
    def ds_geometry(self):
        out = []
        for i in self.f_critical_edges():
            v1, v2 = self.O.edge_verts(self.dims, i)
            for v in (v1, v2):
                pf_v = self.at(self.pv[0], v)
                if pf_v == -1:
                    continue
                if pf_v != self.at(self.pv[1], v):
                    out.append((i, v))
                    break
        return sorted(out)

    # get_false_as_vpath main_3d.cu:4540-4610: the first false cofacet tet per f-critical face
    def as_geometry(self):
        out = []
        for i in self.f_critical_faces():
            for tet in self.face_cofacets(i):
                dv = self.is_false(tet)
                if dv is not None:
                    out.append((tet, dv))
                    break
        return sorted(out)

    # get_false_ds_vpath_connectivity main_3d.cu:4961-5062 with f labels of :4613-4687; the
    # trouble maker is not reset between the two endpoint walks (:4969, :5020-5023)
    def ds_connectivity(self):
        out = []
        for i in self.f_critical_edges():
            v1, v2 = self.O.edge_verts(self.dims, i)
            trouble = -1
            for v0 in (v1, v2):
                lab = self.min_label(v0, 0)
                v = v0
                for _ in range(10 ** 6):
                    dg = self.at(self.pv[1], v)
                    if dg == -1:
                        break
                    if self.at(self.pv[0], v) != dg and trouble == -1:
                        trouble = v
                    a, b = self.O.edge_verts(self.dims, dg)
                    v = b if a == v else a
                if v != lab:
                    out.append((i, trouble))
                    break
        return sorted(out)

    # get_false_as_vpath_connectivity main_3d.cu:4761-4835 over g-critical faces, f labels of
    # get_orginal_saddle_max_pairs :4689-4759 indexed by face id (-1 where f has none)
    def as_connectivity(self):
        fcrit = set(self.f_critical_faces())
        out = []
        for i in [int(x) for x in np.nonzero(self.pf[1] == -1)[0]]:
            cof = self.face_cofacets(i)
            for num, tet in enumerate(cof):
                lab = self.paired_max(tet, 0) if i in fcrit else -1
                if self.paired_max(tet, 1) != lab:
                    dv = self.is_false(tet)
                    if dv is not None:
                        out.append((tet, dv))
                        break
        return sorted(out)


def _cases(oracle):
    out = []
    for seed, shape, rel, kind in [(0, (9, 8, 7), 2e-2, "gm"), (1, (10, 9, 8), 1e-2, "gm"), (2, (8, 10, 9), 3e-2, "fbm"),
                                   (3, (11, 9, 8), 1e-2, "fbm"), (4, (12, 11, 10), 5e-3, "gm")]:
        if kind == "gm":
            f = fields.gaussian_mixture(shape, 4, seed=seed, noise=0.03)
        else:
            f = fields.fbm(shape, seed=seed)
        eps, rng = fields.eps_from_rel(f, rel)
        g0 = fields.quantize(f, eps)
        g1 = oracle.correct(f, g0, eps, rng, oracle.flags())["g"]
        out += [(f, g0), (f, g1)]
    return out


def test_descending_and_ascending_geometry_checks(oracle):
    seen = [0, 0]
    for f, g in _cases(oracle):
        w = Walker(oracle, f, g)
        got = oracle.paths(f, g, oracle.flags(preserve_vpath=1, preserve_geometry=1))
        assert sorted(zip(got["ds"].tolist(), got["ds_rec"].tolist())) == w.ds_geometry()
        assert sorted(zip(got["as_"].tolist(), got["as_rec"].tolist())) == w.as_geometry()
        seen[0] += len(got["ds"])
        seen[1] += len(got["as_"])
    assert seen[0] > 0 and seen[1] > 0


def test_connectivity_checks_and_labels(oracle):
    seen = [0, 0]
    for f, g in _cases(oracle):
        w = Walker(oracle, f, g)
        got = oracle.paths(f, g, oracle.flags(preserve_vpath=1, preserve_geometry=0))
        assert sorted(zip(got["ds"].tolist(), got["ds_rec"].tolist())) == w.ds_connectivity()
        assert sorted(zip(got["as_"].tolist(), got["as_rec"].tolist())) == w.as_connectivity()
        seen[0] += len(got["ds"])
        seen[1] += len(got["as_"])
    assert seen[0] > 0 and seen[1] > 0


@pytest.mark.parametrize("seed", [0, 3])
def test_labels_reach_critical_cells(oracle, seed):
    """Every f label is a critical cell of f reached by a strictly descending (minima) or
    ascending (maxima) walk: labels end where the gradient has no partner (or at the boundary)."""
    f = fields.gaussian_mixture((9, 9, 8), 5, seed=seed, noise=0.03)
    w = Walker(oracle, f, f)
    for i in w.f_critical_edges()[:200]:
        for v in oracle.edge_verts(w.dims, i):
            m = w.min_label(v, 0)
            assert w.pv[0][m] == -1 and f.reshape(-1)[m] <= f.reshape(-1)[v]
    for i in w.f_critical_faces()[:200]:
        for tet in w.face_cofacets(i):
            q = w.paired_max(tet, 0)
            assert q == -1 or w.pt[0][q] == -1
