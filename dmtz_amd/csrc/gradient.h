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

// Out-of-line lower star of vertex v, shared by all pairing lookups (one copy per field type
// instead of one inlined copy per lookup site).
template <class T>
DMTZ_HD_CALL void lower_star_at(const T* fld, Dims d, i64 v, LowerStar* ls) {
    T val[15];
    gather_star(fld, d, v, val);
    lower_star(val, *ls);
}

template <class T>
struct FieldGrad {
    const T* fld;
    Dims d;
    Counts c;

    DMTZ_HD FieldGrad(const T* f, const Dims& dd, const Counts& cc) : fld(f), d(dd), c(cc) {}

    // compare_vertices main_3d.cu:1043-1049: a above b.
    DMTZ_HD bool above(i64 a, i64 b) const {
        T va = fld[a], vb = fld[b];
        return va > vb || (va == vb && a > b);
    }
    DMTZ_HD void star(i64 v, LowerStar& ls) const { lower_star_at<T>(fld, d, v, &ls); }
    // Inlined evaluation for the dense per-vertex kernels.
    DMTZ_HD void star_inline(i64 v, LowerStar& ls) const {
        T val[15];
        gather_star(fld, d, v, val);
        lower_star(val, ls);
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
        LowerStar ls;
        star(v, ls);
        return ls.s < 0 ? -1 : edge_of(v, ls.s);
    }

    // paired_edges[e]
    DMTZ_HD i64 pe(i64 e) const {
        if (e < 0 || e >= c.E) return -1;
        i64 a, b;
        edge_verts(d, c, e, a, b);
        i64 m = above(a, b) ? a : b, o = (m == a) ? b : a;
        int sl = slot_between(d, m, o);
        LowerStar ls;
        star(m, ls);
        if (sl == ls.s) return m;
        int k = ls.e0_of(sl);
        if (k != 15) return tri_id3(d, c, m, o, nb(m, k));
        k = ls.e1_of(sl);
        if (k != 15) return tri_id3(d, c, m, o, nb(m, k));
        return -1;
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
        int s1 = slot_between(d, m, o1), s2 = slot_between(d, m, o2);
        int lt = ST.tri_of[s1][s2];
        LowerStar ls;
        star(m, ls);
        bool hi = (ls.thi >> lt) & 1ull;
        int i = ST.tri_i[lt], j = ST.tri_j[lt];
        int sa = hi ? i : j, sb = hi ? j : i;
        i64 va = nb(m, sa), vb = nb(m, sb);
        if ((ls.tE0 >> lt) & 1ull) return edge_of(m, sa);
        if ((ls.tT0 >> lt) & 1ull) return tet_id4(d, c, m, va, vb, nb(m, ST.tri_k[lt][(ls.tD0 >> lt) & 1ull]));
        if ((ls.tE1 >> lt) & 1ull) return edge_of(m, sb);
        if ((ls.tT1 >> lt) & 1ull) return tet_id4(d, c, m, va, vb, nb(m, ST.tri_k[lt][(ls.tD1 >> lt) & 1ull]));
        return -1;
    }

    // paired_tetrahedras[q]
    DMTZ_HD i64 pt(i64 q) const {
        if (q < 0 || q >= c.T) return -1;
        i64 w[4];
        tet_verts(d, q, w[0], w[1], w[2], w[3]);
        int mi = 0;
        for (int k = 1; k < 4; k++)
            if (above(w[k], w[mi])) mi = k;
        i64 m = w[mi];
        i64 o[3];
        int n = 0;
        for (int k = 0; k < 4; k++)
            if (k != mi) o[n++] = w[k];
        // order the other three: p1 above p2 above p3
        if (above(o[1], o[0])) { i64 x = o[0]; o[0] = o[1]; o[1] = x; }
        if (above(o[2], o[0])) { i64 x = o[0]; o[0] = o[2]; o[2] = x; }
        if (above(o[2], o[1])) { i64 x = o[1]; o[1] = o[2]; o[2] = x; }
        int s1 = slot_between(d, m, o[0]), s2 = slot_between(d, m, o[1]), s3 = slot_between(d, m, o[2]);
        uint32_t key = (1u << s1) | (1u << s2) | (1u << s3);
        int lq = 0;
        for (int k = 0; k < 24; k++) {
            uint32_t kk = (1u << ST.tet_s[k][0]) | (1u << ST.tet_s[k][1]) | (1u << ST.tet_s[k][2]);
            lq = (kk == key) ? k : lq;
        }
        LowerStar ls;
        star(m, ls);
        if ((ls.qT0 >> lq) & 1u) return tri_id3(d, c, m, o[0], o[1]);
        if ((ls.qT1 >> lq) & 1u) return tri_id3(d, c, m, o[0], o[2]);
        return -1;
    }
};

}  // namespace dmtz
