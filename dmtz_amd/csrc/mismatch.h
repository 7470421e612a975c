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
// at w. (The device gathers instead, crit_gather below; the host check compares all three.)
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

// The gather form of cell_at_other_vertices. For a vertex u and its neighbour w in slot s
// (w = u + DIR[s]): the local cells of u that contain w (cont_lo[s]: bits below LC_TET,
// cont_hi[s]: tet bits) and the local index of each of them in w's star (at_w[s][lc], -1 for a
// cell without w).
struct SpreadTables {
    uint64_t cont_lo[14];
    uint32_t cont_hi[14];
    int8_t at_w[14][LC_COUNT];
};

namespace detail {
constexpr int c_slot_from(int from, int to) {
    const int dx = DIR[to][0] - DIR[from][0], dy = DIR[to][1] - DIR[from][1], dz = DIR[to][2] - DIR[from][2];
    return (dx < -1 || dx > 1 || dy < -1 || dy > 1 || dz < -1 || dz > 1)
               ? -1
               : ST.slot_of[(dx + 1) * 9 + (dy + 1) * 3 + (dz + 1)];
}
constexpr int c_tet_of(int a, int b, int c) {
    const uint32_t key = (1u << a) | (1u << b) | (1u << c);
    int r = -1;
    for (int k = 0; k < 24; k++)
        if (((1u << ST.tet_s[k][0]) | (1u << ST.tet_s[k][1]) | (1u << ST.tet_s[k][2])) == key) r = k;
    return r;
}
constexpr void c_put(SpreadTables& t, int s, int lc, int at) {
    if (lc < LC_TET) t.cont_lo[s] |= 1ull << lc;
    else t.cont_hi[s] |= 1u << (lc - LC_TET);
    t.at_w[s][lc] = (int8_t)at;
}
constexpr SpreadTables make_spread_tables() {
    SpreadTables t{};
    for (int s = 0; s < 14; s++) {
        t.cont_lo[s] = 0;
        t.cont_hi[s] = 0;
        for (int lc = 0; lc < LC_COUNT; lc++) t.at_w[s][lc] = -1;
    }
    for (int e = 0; e < 14; e++) c_put(t, e, LC_EDGE + e, LC_EDGE + (e ^ 1));
    for (int tr = 0; tr < 36; tr++) {
        const int i = ST.tri_i[tr], j = ST.tri_j[tr];
        c_put(t, i, LC_TRI + tr, LC_TRI + ST.tri_of[i ^ 1][c_slot_from(i, j)]);
        c_put(t, j, LC_TRI + tr, LC_TRI + ST.tri_of[j ^ 1][c_slot_from(j, i)]);
    }
    for (int q = 0; q < 24; q++) {
        const int s0 = ST.tet_s[q][0], s1 = ST.tet_s[q][1], s2 = ST.tet_s[q][2];
        c_put(t, s0, LC_TET + q, LC_TET + c_tet_of(s0 ^ 1, c_slot_from(s0, s1), c_slot_from(s0, s2)));
        c_put(t, s1, LC_TET + q, LC_TET + c_tet_of(s1 ^ 1, c_slot_from(s1, s0), c_slot_from(s1, s2)));
        c_put(t, s2, LC_TET + q, LC_TET + c_tet_of(s2 ^ 1, c_slot_from(s2, s0), c_slot_from(s2, s1)));
    }
    return t;
}
}  // namespace detail

constexpr SpreadTables SP = detail::make_spread_tables();

// The critical status in f of every local cell of v, gathered: own(u, lo, hi) gives the critical
// cells vertex u owns (code_masks of its code, restricted to its members). v's own bits, plus, for
// each neighbour u = v + DIR[s] in the grid, u's owned critical cells that contain v, moved to
// their local indices at v (v is in slot s ^ 1 of u's star). Every cell of v's star is owned by
// v or by one of its 14 neighbours, so this equals local_crit and crit_scatter_spread, with
// plain reads of the neighbours' masks instead of lookups or atomics.
template <class Own>
DMTZ_HD void crit_gather(const Dims& d, i64 v, Own own, uint64_t& lo, uint32_t& hi) {
    own(v, lo, hi);
    i64 x, y, z;
    vxyz(d, v, x, y, z);
    for (int s = 0; s < 14; s++) {
        const i64 nx = x + DIR[s][0], ny = y + DIR[s][1], nz = z + DIR[s][2];
        if (!inb(d, nx, ny, nz)) continue;
        uint64_t ulo;
        uint32_t uhi;
        own(vidx(d, nx, ny, nz), ulo, uhi);
        const int r = s ^ 1;
        uint64_t m = ulo & SP.cont_lo[r];
        uint32_t h = uhi & SP.cont_hi[r];
        while (m | h) {
            const int c = m ? __builtin_ctzll(m) : LC_TET + __builtin_ctz(h);
            if (m) m &= m - 1;
            else h &= h - 1;
            const int w = SP.at_w[r][c];
            if (w < LC_TET) lo |= 1ull << w;
            else hi |= 1u << (w - LC_TET);
        }
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
