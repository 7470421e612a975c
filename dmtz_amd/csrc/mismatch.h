// Mismatch detection (get_false_minimum / _saddle / _2_saddle / _maximum, main_3d.cu:2492-2550)
// from per-vertex lower stars.
//
// The reference tests crit_f(cell) XOR crit_g(cell) for every simplex. A cell's status in a field
// is decided in the lower star of its highest vertex in that field (its owner), and the owners in
// f and g can differ. f does not change during the correction, so crit_f is evaluated once for
// every local cell of every vertex (local_crit); each sweep then tests every cell at its owner in
// g from the g star code alone (owner_mismatch), with no lookups.
#pragma once
#include "gradient.h"

namespace dmtz {

// Membership of the 75 local cells in a lower star (same split as crit_mask).
DMTZ_HD void member_mask(const LowerStar& r, uint64_t& lo, uint32_t& hi) {
    lo = 1ull | ((uint64_t)r.L << 1) | (r.tin << 15);
    hi = r.qin;
}

DMTZ_HD bool lc_bit(uint64_t lo, uint32_t hi, int lc) {
    return lc < 51 ? ((lo >> lc) & 1ull) != 0 : ((hi >> (lc - 51)) & 1u) != 0;
}

DMTZ_HD int lc_type(int lc) { return lc == LC_VERT ? 0 : (lc < LC_TRI ? 1 : (lc < LC_TET ? 2 : 3)); }

// Reference simplex id of local cell lc of vertex v.
DMTZ_HD i64 local_cell_id(const Dims& d, const Counts& c, i64 v, int lc) {
    i64 x, y, z;
    vxyz(d, v, x, y, z);
    auto nb = [&](int k) { return vidx(d, x + DIR[k][0], y + DIR[k][1], z + DIR[k][2]); };
    if (lc == LC_VERT) return v;
    if (lc < LC_TRI) return edge_id_dir(d, c, x, y, z, lc - LC_EDGE);
    if (lc < LC_TET) {
        int t = lc - LC_TRI;
        return tri_id3(d, c, v, nb(ST.tri_i[t]), nb(ST.tri_j[t]));
    }
    int q = lc - LC_TET;
    return tet_id4(d, c, v, nb(ST.tet_s[q][0]), nb(ST.tet_s[q][1]), nb(ST.tet_s[q][2]));
}

template <class T>
DMTZ_HD bool is_crit_lookup(const FieldGrad<T>& X, int type, i64 id) {
    switch (type) {
        case 0: return X.pv(id) == -1;
        case 1: return X.pe(id) == -1;
        case 2: return X.pf(id) == -1;
        default: return X.pt(id) == -1;
    }
}

// Existing local cells of vertex v (every cell whose vertices lie in the grid), as (lo, hi).
DMTZ_HD void local_cells_exist(const Dims& d, i64 v, uint64_t& lo, uint32_t& hi) {
    i64 x, y, z;
    vxyz(d, v, x, y, z);
    if (x > 0 && y > 0 && z > 0 && x < d.W - 1 && y < d.H - 1 && z < d.D - 1) {
        lo = ~0ull >> (64 - LC_TET);
        hi = (1u << (LC_COUNT - LC_TET)) - 1u;
        return;
    }
    uint32_t s = 0;  // neighbour slots inside the grid
    for (int k = 0; k < 14; k++)
        if (inb(d, x + DIR[k][0], y + DIR[k][1], z + DIR[k][2])) s |= 1u << k;
    lo = 1ull | ((uint64_t)s << LC_EDGE);
    for (int t = 0; t < 36; t++)
        if ((s >> ST.tri_i[t] & 1u) && (s >> ST.tri_j[t] & 1u)) lo |= 1ull << (LC_TRI + t);
    hi = 0;
    for (int q = 0; q < 24; q++)
        if ((s >> ST.tet_s[q][0] & 1u) && (s >> ST.tet_s[q][1] & 1u) && (s >> ST.tet_s[q][2] & 1u)) hi |= 1u << q;
}

// Critical status in field F of every existing local cell of v: the bits of the cells v owns in
// F come from v's star code, the others from a pairing lookup at their owner. F is fixed during
// the correction, so this is evaluated once per run for f.
template <class T>
DMTZ_HD void local_crit(const FieldGrad<T>& F, i64 v, uint64_t& lo, uint32_t& hi) {
    uint64_t c_lo, i_lo, e_lo;
    uint32_t c_hi, i_hi, e_hi;
    code_masks(F.code[v], c_lo, c_hi, i_lo, i_hi);
    local_cells_exist(F.d, v, e_lo, e_hi);
    lo = c_lo & i_lo;
    hi = c_hi & i_hi;
    uint64_t x_lo = e_lo & ~i_lo;
    uint32_t x_hi = e_hi & ~i_hi;
    while (x_lo | x_hi) {
        int lc = x_lo ? __builtin_ctzll(x_lo) : 51 + __builtin_ctz(x_hi);
        if (x_lo) x_lo &= x_lo - 1;
        else x_hi &= x_hi - 1;
        if (is_crit_lookup(F, lc_type(lc), local_cell_id(F.d, F.c, v, lc))) {
            if (lc < 51) lo |= 1ull << lc;
            else hi |= 1u << (lc - 51);
        }
    }
}

// Slot of a stencil offset, -1 when the offset is not one of the 14 directions.
DMTZ_HD int offset_slot(int dx, int dy, int dz) { return ST.slot_of[(dx + 1) * 9 + (dy + 1) * 3 + (dz + 1)]; }

// Slot, seen from the vertex in direction `from`, of the vertex in direction `to` (both slots of
// the same star; the two vertices are adjacent whenever they share a cell with the centre).
DMTZ_HD int slot_from(int from, int to) {
    return offset_slot(DIR[to][0] - DIR[from][0], DIR[to][1] - DIR[from][1], DIR[to][2] - DIR[from][2]);
}

// Every vertex w != u of local cell lc (> 0) of vertex u, with the cell's local index at w:
// fn(w, lc_at_w). Slot s ^ 1 is the direction opposite to slot s.
template <class Fn>
DMTZ_HD void cell_at_other_vertices(const Dims& d, i64 u, int lc, Fn fn) {
    i64 x, y, z;
    vxyz(d, u, x, y, z);
    auto at = [&](int s) { return vidx(d, x + DIR[s][0], y + DIR[s][1], z + DIR[s][2]); };
    if (lc < LC_TRI) {
        const int s = lc - LC_EDGE;
        fn(at(s), LC_EDGE + (s ^ 1));
    } else if (lc < LC_TET) {
        const int t = lc - LC_TRI, i = ST.tri_i[t], j = ST.tri_j[t];
        fn(at(i), LC_TRI + ST.tri_of[i ^ 1][slot_from(i, j)]);
        fn(at(j), LC_TRI + ST.tri_of[j ^ 1][slot_from(j, i)]);
    } else {
        const int q = lc - LC_TET;
        const int s0 = ST.tet_s[q][0], s1 = ST.tet_s[q][1], s2 = ST.tet_s[q][2];
        auto tet_at = [&](int a, int b, int c) {  // tet of the centre's neighbours in slots a, b, c
            const uint32_t key = (1u << a) | (1u << b) | (1u << c);
            int r = 0;
            for (int k = 0; k < 24; k++) {
                const uint32_t kk = (1u << ST.tet_s[k][0]) | (1u << ST.tet_s[k][1]) | (1u << ST.tet_s[k][2]);
                r = kk == key ? k : r;
            }
            return LC_TET + r;
        };
        fn(at(s0), tet_at(s0 ^ 1, slot_from(s0, s1), slot_from(s0, s2)));
        fn(at(s1), tet_at(s1 ^ 1, slot_from(s1, s0), slot_from(s1, s2)));
        fn(at(s2), tet_at(s2 ^ 1, slot_from(s2, s0), slot_from(s2, s1)));
    }
}

// The critical status in f of every local cell, written the other way round from local_crit: a
// vertex's mask first holds the critical cells it owns (code_masks of its code); then every
// vertex spreads its owned critical cells to their other vertices: spread(w, lc) adds cell lc
// at w.
template <class Spread>
DMTZ_HD void crit_scatter_spread(const Dims& d, const StarCode& code, i64 v, Spread spread) {
    uint64_t c_lo, i_lo;
    uint32_t c_hi, i_hi;
    code_masks(code, c_lo, c_hi, i_lo, i_hi);
    uint64_t m = c_lo & i_lo & ~1ull;
    uint32_t h = c_hi & i_hi;
    while (m) {
        const int lc = __builtin_ctzll(m);
        m &= m - 1;
        cell_at_other_vertices(d, v, lc, spread);
    }
    while (h) {
        const int lc = LC_TET + __builtin_ctz(h);
        h &= h - 1;
        cell_at_other_vertices(d, v, lc, spread);
    }
}

// Mismatched cells owned by v in g (main_3d.cu:2492-2550: crit_f XOR crit_g per cell), from v's
// g star code and the f-critical mask of v's local cells. Every cell has exactly one owner in g,
// so the union over all vertices lists each mismatched cell once.
DMTZ_HD void owner_mismatch(const StarCode& cg, uint64_t cf_lo, uint32_t cf_hi, uint64_t& mlo, uint32_t& mhi) {
    uint64_t c_lo, i_lo;
    uint32_t c_hi, i_hi;
    code_masks(cg, c_lo, c_hi, i_lo, i_hi);
    mlo = (c_lo ^ cf_lo) & i_lo;
    mhi = (c_hi ^ cf_hi) & i_hi;
}

}  // namespace dmtz
