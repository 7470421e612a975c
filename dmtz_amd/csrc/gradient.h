// The reference's pairing arrays answered from per-vertex lower stars.
//
// The reference stores paired_vertices / paired_edges / paired_faces / paired_tetrahedras for
// f and g as dense int32 arrays (main_3d.cu:76-101, allocated main_3d.cu:6474-6517) and reads
// them with overloaded meanings: paired_edges[e] holds a vertex id or a face id, paired_faces[t]
// an edge id or a tet id, and the repair rules read them at IDs of another type
// (e.g. main_3d.cu:2719, 3124, 3522). Here nothing is stored: a read of array X at id i decodes
// i as a simplex of X's dimension, finds that simplex's highest vertex m, evaluates the lower
// star of m and returns the partner's reference id. The reads are total over each id range;
// an id outside [0, size) reads as -1 ("unpaired"), the resolution of the reference's
// out-of-range reads.
#pragma once
#include "lower_star.h"

namespace dmtz {

template <class T>
DMTZ_HD T pos_inf();
template <>
DMTZ_HD float pos_inf<float>() { return __builtin_huge_valf(); }
template <>
DMTZ_HD double pos_inf<double>() { return __builtin_huge_val(); }

// Closed-star values of vertex v (val[14] = v), +inf for off-grid neighbours.
template <class T>
DMTZ_HD void gather_star(const T* fld, const Dims& d, i64 v, T* val) {
    i64 x, y, z;
    vxyz(d, v, x, y, z);
#pragma unroll
    for (int k = 0; k < 14; k++) {
        i64 nx = x + DIR[k][0], ny = y + DIR[k][1], nz = z + DIR[k][2];
        val[k] = inb(d, nx, ny, nz) ? fld[vidx(d, nx, ny, nz)] : pos_inf<T>();
    }
    val[14] = fld[v];
}

// Lower star of vertex v from the field values (the dense classification step).
template <class T>
DMTZ_HD void star_of(const T* fld, const Dims& d, i64 v, LowerStar& ls) {
    T val[15];
    gather_star(fld, d, v, val);
    lower_star(val, ls);
}

// Pairing lookups over precomputed star codes (code[v] = encode_star(lower star of v)). Finding
// a cell's highest vertex needs only the cell's own values; the partner then comes from that
// vertex's 32-byte code.
template <class T>
struct FieldGrad {
    const T* fld;
    const StarCode* code;
    Dims d;
    Counts c;

    DMTZ_HD FieldGrad(const T* f, const StarCode* cd, const Dims& dd, const Counts& cc)
        : fld(f), code(cd), d(dd), c(cc) {}

    // compare_vertices main_3d.cu:1043-1049: a above b.
    DMTZ_HD bool above(i64 a, i64 b) const {
        T va = fld[a], vb = fld[b];
        return va > vb || (va == vb && a > b);
    }
    DMTZ_HD i64 nb(i64 v, int slot) const {
        i64 x, y, z;
        vxyz(d, v, x, y, z);
        return vidx(d, x + DIR[slot][0], y + DIR[slot][1], z + DIR[slot][2]);
    }
    DMTZ_HD i64 edge_of(i64 v, int slot) const {
        i64 x, y, z;
        vxyz(d, v, x, y, z);
        return edge_id_dir(d, c, x, y, z, slot);
    }

    // paired_vertices[v]
    DMTZ_HD i64 pv(i64 v) const {
        if (v < 0 || v >= c.N) return -1;
        int s = code_s(code[v]);
        return s == 15 ? -1 : edge_of(v, s);
    }

    // paired_edges[e]
    DMTZ_HD i64 pe(i64 e) const {
        if (e < 0 || e >= c.E) return -1;
        i64 a, b;
        edge_verts(d, c, e, a, b);
        i64 m = above(a, b) ? a : b, o = (m == a) ? b : a;
        int sl = slot_between(d, m, o);
        int nib = code_edge(code[m], sl);
        if (nib == 14) return m;
        if (nib == 15) return -1;
        return tri_id3(d, c, m, o, nb(m, nib));
    }

    // paired_faces[t]
    DMTZ_HD i64 pf(i64 t) const {
        if (t < 0 || t >= c.F) return -1;
        i64 p, q, r;
        tri_verts(d, c, t, p, q, r);
        i64 m = p;
        if (above(q, m)) m = q;
        if (above(r, m)) m = r;
        i64 o1 = (m == p) ? q : p;
        i64 o2 = (m == r) ? q : r;
        int lt = ST.tri_of[slot_between(d, m, o1)][slot_between(d, m, o2)];
        int cc = code_tri(code[m], lt);
        int i = ST.tri_i[lt], j = ST.tri_j[lt];
        switch (cc) {
            case 0: return edge_of(m, i);
            case 1: return edge_of(m, j);
            case 2: return tet_id4(d, c, m, nb(m, i), nb(m, j), nb(m, ST.tri_k[lt][0]));
            case 3: return tet_id4(d, c, m, nb(m, i), nb(m, j), nb(m, ST.tri_k[lt][1]));
            default: return -1;
        }
    }

    // paired_tetrahedras[q]
    DMTZ_HD i64 pt(i64 q) const {
        if (q < 0 || q >= c.T) return -1;
        i64 w[4];
        tet_verts(d, q, w[0], w[1], w[2], w[3]);
        i64 m = w[0];
        for (int k = 1; k < 4; k++)
            if (above(w[k], m)) m = w[k];
        uint32_t key = 0;
        for (int k = 0; k < 4; k++)
            if (w[k] != m) key |= 1u << slot_between(d, m, w[k]);
        int lq = 0;
        for (int k = 0; k < 24; k++) {
            uint32_t kk = (1u << ST.tet_s[k][0]) | (1u << ST.tet_s[k][1]) | (1u << ST.tet_s[k][2]);
            lq = (kk == key) ? k : lq;
        }
        int cc = code_tet(code[m], lq);
        if (cc == 3) return -1;
        int s0 = ST.tet_s[lq][cc == 0 ? 1 : 0];
        int s1 = ST.tet_s[lq][cc == 2 ? 1 : 2];
        return tri_id3(d, c, m, nb(m, s0), nb(m, s1));
    }
};

}  // namespace dmtz
