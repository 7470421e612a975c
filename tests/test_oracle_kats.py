"""Hand-worked known answers for every top-level branch of the four critical-cell repair rules.

Each case is a tiny integer field f and a g that differs from f at one or two vertices. The
expected edit set of one mismatched cell was derived by hand by reading the cited reference lines
with these numbers (the derivation is in the comments); the oracle's restatement of that rule
must produce exactly that set. Simplex IDs come from the oracle's ID maps, which
test_oracle.py pins by closed-form counts, bijections and the clique structure; the pairing facts
a derivation starts from are checked against the oracle's gradient, pinned there by the literal
14^k search, the Euler characteristic and matching validity.

Grids are x fastest: vertex = x + W*y + W*H*z. Values are listed in vertex order.
"""
import pytest

from corpus import kat_field as field


def cell_id(oracle, shape, verts):
    D, H, W = shape
    dims = (W, H, D)
    if len(verts) == 1:
        return verts[0]
    return (oracle.edge_id, oracle.tri_id, oracle.tet_id)[len(verts) - 2](dims, *verts)


# (branch, input, cell vertices, expected edit set). The derivations:
KATS = [
    # fix_minimum :2560-2576. Vertex 8 = (2,2,0), f = 17; its neighbours 7, 5, 17, 16 (f = 10, 8,
    # 14, 12) are all lower, steepest 5, so f pairs 8 with edge (8,5) (processVertices
    # :1930-1988). g(8) = 0.5 is below them all: a g minimum, so 8 is a false minimum.
    # paired_vertices[8] != -1; getVertexIDsFromEdgeID gives (8, 5); vertex == id, so v2 = 5 is
    # edited.
    ("MIN_F_PAIRED", "A", (8,), [5]),
    # fix_minimum :2577-2585. Vertex 6 = (0,0,1): neighbours 7, 9, 0, 1 (f = 7, 10, 9, 11) are all
    # higher, so 6 is a minimum of f; g(1) = 2.5 gives it a lower neighbour in g. f leaves it
    # unpaired: the rule edits 6 itself.
    ("MIN_F_CRITICAL", "B", (6,), [6]),
    # fix_saddle :2602-2641. Edge (7,6): vertex 7 = (1,2,0) has lower neighbours 6, 4, 5, 15
    # (f = 1, 5, 8, 3), steepest 6, so f pairs 7 with this edge. In g, 8 (g = 0.5) becomes 7's
    # steepest, so the edge is left critical. v1, v2 = (7, 6); v2 != paired vertex 7, so v1 = 7,
    # v2 = 6. dec_paired_vertices[7] = edge (8,7): false_v2 = 8 (:2619-2623). g(6) = 1 > g(8) =
    # 0.5 (:2626), so 6 is edited.
    ("S1_FP_VERTEX", "A", (7, 6), [6]),
    # fix_saddle :2643-2827. Edge (4,1), f-paired with triangle (4,2,1); third vertex v3 = 2
    # (:2657-2658). v1_edge = g(4) > g(2): 1.5 > 4 no; v2_edge = g(1) > g(2): yes (:2676-2677), so
    # :2684 fails. o_v1 = f(4) > f(2): 5 > 4 yes; o_v2 = 11 > 4 yes (:2680-2681): :2700 holds and
    # v3 = 2 is edited.
    ("S1_FP_FACE", "C", (4, 1), [2]),
    # fix_saddle :2839-2850. Edge (14,2): in g vertex 2 = (2,0,0) (g = 11.5) has lower neighbours
    # 5, 10, 4, 14 (g = 11, 9, 4.5, 3), steepest 14, so g pairs 2 with this edge; in f, 2's
    # steepest is 4 (f = 2) and the edge is critical. dec_paired_edges[id] = 2 is a vertex paired
    # back: v_id = 2 is edited.
    ("S1_FN_VERTEX", "D", (14, 2), [2]),
    # fix_saddle :2852-3035. Edge (2,1), g-paired with triangle (4,2,1); v1, v2 = (2, 1), v3 = 4
    # (:2861). v1_edge = g(1) > g(4): 11 > 1.5 yes; v2_edge = g(2) > g(4): 4 > 1.5 yes; or_v1 =
    # f(1) > f(4): 11 > 5 yes; or_v2 = f(2) > f(4): 4 > 5 no (:2886-2889). v1_edge == v2_edge
    # (:2891): v1_edge == or_v1, so v2 is not edited; v2_edge != or_v2, so v1 = 2 is (:2903).
    ("S1_FN_FACE", "C", (2, 1), [2]),
    # fix_2saddle :3057-3223. Triangle (7,5,4), f-paired with edge (7,5); v1, v2 = (7, 5), v3 = 4.
    # v1_edge = g(5) > g(4): 8 > 5, v2_edge = g(7) > g(4): 10 > 5, both hold (:3097). f_id =
    # paired_edges[edge] is this triangle and pairs back (:3099-3100); its vertex off the edge,
    # 4, is edited.
    ("S2_FP_EDGE", "A", (7, 5, 4), [4]),
    # fix_2saddle :3226-3477. Triangle (16,15,13), f-paired with tet (16,15,13,4); v4 = 4
    # (:3233-3235). or_f1 = f(13) > f(4): 13 > 2, or_f2 = f(15) > f(4): 4 > 2, or_f3 = f(16) >
    # f(4): 8 > 2 all hold, so :3260 fails and the else at :3351 edits v4 = 4.
    ("S2_FP_TET", "D", (16, 15, 13), [4]),
    # fix_2saddle :3488-3610. Triangle (16,14,13) is critical in f; g pairs it with edge (16,13).
    # sortedV1, sortedV2 = 13, 16 (g(16) = 8 < g(13) = 13). In f, vertex 13's steepest is 4
    # (f = 2), so the edge is not vertex-paired, and in pass E0 (:2057-2071) its lowest common
    # neighbour below both ends is 4 (candidates 15, 14, 4 with f = 4, 3, 2): f pairs the edge
    # with triangle (16,13,4). true_face pairs back (:3522); its vertex other than 13 and 16,
    # 4, is edited.
    ("S2_FN_EDGE", "D", (16, 14, 13), [4]),
    # fix_2saddle :3613-3980. Triangle (9,6,0) is critical in f; g pairs it with tet (9,6,1,0)
    # (g(1) = 2.5). v1, v2, v3 = 9, 6, 0 and v4 = 1 (:3619-3621). f1_face = g(0) > g(1): 9 > 2.5,
    # f2_face = g(6) > g(1), f3_face = g(9) > g(1): all hold (:3714); or_f1 = f(0) > f(1): 9 > 11
    # fails, so f1_face != or_f1 (:3716) and, f1_face being set, v3 = 0 is edited (:3719-3725).
    ("S2_FN_TET", "B", (9, 6, 0), [0]),
    # fix_maximum :3995-4097. Tet (16,7,5,4), f-paired with triangle (16,7,5); v4 = 4. f(16) = 12,
    # f(7) = 10, f(5) = 8 are all above f(4) = 5, so :4028 holds; delta(4) > the buffer's
    # sentinel (:4032), and 4 is edited.
    ("MAX_FP", "A", (16, 7, 5, 4), [4]),
    # fix_maximum :4098-4298. Tet (16,14,13,5) is critical in f; g pairs it with triangle
    # (16,13,5); v4 = 14. g(16) = 8, g(13) = 13, g(5) = 11 are above g(14) = 3 (:4169), and in f
    # too (8, 13, 11 > 3), so none of :4171-4200 edits and :4202 reads f's partner of the
    # triangle: its top edge (13,5) takes triangle (13,5,4) in pass E0, so the triangle is still
    # free in pass T0, where its lowest common neighbour below all three is 4 (candidates 14, 4
    # with f = 3, 2): tet (16,13,5,4). Its vertex off the triangle, 4, is edited (:4206-4219).
    ("MAX_FN", "D", (16, 14, 13, 5), [4]),
]


@pytest.mark.parametrize("branch,inp,verts,want", KATS, ids=[k[0] for k in KATS])
def test_repair_rule_known_answer(oracle, branch, inp, verts, want):
    f, g = field(inp)
    cell = cell_id(oracle, f.shape, verts)
    types = {1: 0, 2: 1, 3: 2, 4: 3}
    got = oracle.fix_one(f, g, 100.0, 100.0, types[len(verts)], cell)
    assert got is not None, "the cell is not mismatched"
    assert got == want, (branch, got, want)


def test_premises_of_the_derivations(oracle):
    """The pairings the derivations start from, read from the gradients."""
    D, H, W = (2, 3, 3)
    dims = (W, H, D)
    f, g = field("A")
    pvf = oracle.gradient(f)[0]
    pvg = oracle.gradient(g)[0]
    assert pvf[8] == oracle.edge_id(dims, 8, 5) and pvg[8] == -1
    assert pvf[7] == oracle.edge_id(dims, 7, 6) and pvg[7] == oracle.edge_id(dims, 8, 7)
    f, g = field("D")
    pvf, pef, pff, _ = oracle.gradient(f)
    pvg = oracle.gradient(g)[0]
    assert pvg[2] == oracle.edge_id(dims, 14, 2) and pvf[2] == oracle.edge_id(dims, 4, 2)
    assert pvf[13] == oracle.edge_id(dims, 13, 4)
    e = oracle.edge_id(dims, 16, 13)
    assert pef[e] == oracle.tri_id(dims, 16, 13, 4)
    t = oracle.tri_id(dims, 16, 13, 5)
    assert pff[t] == oracle.tet_id(dims, 16, 13, 5, 4)
    assert pef[oracle.edge_id(dims, 13, 5)] == oracle.tri_id(dims, 13, 5, 4)
