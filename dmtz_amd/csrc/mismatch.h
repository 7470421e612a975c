// Mismatch detection (get_false_minimum / _saddle / _2_saddle / _maximum, main_3d.cu:2492-2550)
// from per-vertex lower stars.
//
// The reference tests crit_f(cell) XOR crit_g(cell) for every simplex. A cell's status in a field
// is decided in the lower star of its highest vertex in that field (its owner), and the owners in
// f and g can differ. f does not change during the correction, so crit_f is evaluated once for
// every local cell of every vertex (local_crit); each sweep then judges every cell at its owner in
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
