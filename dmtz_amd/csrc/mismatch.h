// Mismatch detection (get_false_minimum / _saddle / _2_saddle / _maximum, main_3d.cu:2492-2550)
// from per-vertex lower stars.
//
// The reference tests crit_f(cell) XOR crit_g(cell) for every simplex. A cell's status in a field
// is decided in the lower star of its highest vertex in that field, and that owner can differ
// between f and g. Per vertex v:
//   * cells owned by v in both fields: XOR of the two local critical bits;
//   * cells owned by v in f only and critical in f: mismatched unless g's pairing lookup
//     (evaluated at the g-owner) says critical;
//   * cells owned by v in g only and critical in g: mismatched unless f's lookup says critical.
// A mismatched cell is thus reported exactly once: by its common owner, or by the owner in the
// field where it is critical.
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

// Mismatched local cells of vertex v: (lo, hi) bit masks over the 75 local cells, from the
// star codes of v in both fields.
template <class T>
DMTZ_HD void vertex_mismatch(const FieldGrad<T>& F, const FieldGrad<T>& G, i64 v, uint64_t& mlo, uint32_t& mhi) {
    uint64_t cf_lo, cg_lo, if_lo, ig_lo;
    uint32_t cf_hi, cg_hi, if_hi, ig_hi;
    code_masks(F.code[v], cf_lo, cf_hi, if_lo, if_hi);
    code_masks(G.code[v], cg_lo, cg_hi, ig_lo, ig_hi);
    mlo = (cf_lo ^ cg_lo) & if_lo & ig_lo;
    mhi = (cf_hi ^ cg_hi) & if_hi & ig_hi;
    uint64_t xf_lo = cf_lo & ~ig_lo, xg_lo = cg_lo & ~if_lo;
    uint32_t xf_hi = cf_hi & ~ig_hi, xg_hi = cg_hi & ~if_hi;
    // cross-owner cells, resolved by a lookup in the other field
    while (xf_lo | xf_hi) {
        int lc = xf_lo ? __builtin_ctzll(xf_lo) : 51 + __builtin_ctz(xf_hi);
        if (xf_lo) xf_lo &= xf_lo - 1;
        else xf_hi &= xf_hi - 1;
        if (!is_crit_lookup(G, lc_type(lc), local_cell_id(F.d, F.c, v, lc))) {
            if (lc < 51) mlo |= 1ull << lc;
            else mhi |= 1u << (lc - 51);
        }
    }
    while (xg_lo | xg_hi) {
        int lc = xg_lo ? __builtin_ctzll(xg_lo) : 51 + __builtin_ctz(xg_hi);
        if (xg_lo) xg_lo &= xg_lo - 1;
        else xg_hi &= xg_hi - 1;
        if (!is_crit_lookup(F, lc_type(lc), local_cell_id(F.d, F.c, v, lc))) {
            if (lc < 51) mlo |= 1ull << lc;
            else mhi |= 1u << (lc - 51);
        }
    }
}

}  // namespace dmtz
