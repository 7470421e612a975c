// Repair rules and separatrix checks of the reference, over the lower-star pairing lookups.
//
// Every rule ends in "add vertex w to this sweep's edit set". In the reference that is
// swap(w, get_delta(w)) into d_deltaBuffer (main_3d.cu:1088-1120); get_delta(w) depends on w
// alone within a sweep, so the buffer after the repair phase is exactly {w -> delta(w)} over the
// edit set, independent of thread order. Here the sink records w; the apply step recomputes
// delta(w). Rules read f and g orders and the four pairing arrays of both fields; the branch
// structure follows the cited lines, including their quirks (noted inline).
#pragma once
#include "edit.h"
#include "gradient.h"

namespace dmtz {

template <class T, class Sink>
struct Repair {
    FieldGrad<T> F;  // original field f ("type 0")
    FieldGrad<T> G;  // decompressed field g ("type 1")
    const uint8_t* counter;
    double eps, range;
    int edit_type;
    Sink sink;

    DMTZ_HD Repair(const T* f, const StarCode* cf, const T* g, const StarCode* cg, const Dims& d, const Counts& c,
                   const uint8_t* cnt, double e, double r, int et, Sink s)
        : F(f, cf, d, c), G(g, cg, d, c), counter(cnt), eps(e), range(r), edit_type(et), sink(s) {}

    DMTZ_HD i64 N() const { return F.c.N; }
    DMTZ_HD void tgt(i64 v) {
        if (v >= 0 && v < N()) sink(v);
    }
    // get_delta main_3d.cu:239-244
    DMTZ_HD double delta(i64 v) const {
        if (edit_type == 1) return halving_delta(eps, counter[v]);
        return EditArith<T>::delta0(F.fld[v], G.fld[v], eps);
    }
    DMTZ_HD bool fa(i64 a, i64 b) const { return F.above(a, b); }
    DMTZ_HD bool ga(i64 a, i64 b) const { return G.above(a, b); }

    static DMTZ_HD i64 tet_other(i64 q1, i64 q2, i64 q3, i64 q4, i64 a, i64 b, i64 c) {
        // first tet vertex not on the face, else the fourth (main_3d.cu:3237-3239)
        i64 v4 = q4;
        if (q1 != a && q1 != b && q1 != c) v4 = q1;
        else if (q2 != a && q2 != b && q2 != c) v4 = q2;
        else if (q3 != a && q3 != b && q3 != c) v4 = q3;
        return v4;
    }
    DMTZ_HD void tri_third(i64 t, i64 v1, i64 v2, i64& v3) const {
        // face t's vertex not on edge (v1, v2), else-if form (main_3d.cu:2657-2658)
        i64 a, b;
        tri_verts(F.d, F.c, t, a, b, v3);
        if (a != v1 && a != v2) v3 = a;
        else if (b != v1 && b != v2) v3 = b;
    }

    // Lowest f-vertex among the common neighbours of (v1, v2) strictly between them in f
    // (main_3d.cu:2942-3003), -1 if none.
    DMTZ_HD i64 lowest_between_f(i64 v1, i64 v2) const {
        int sl = slot_between(F.d, v1, v2);
        if (sl < 0) return -1;
        i64 x, y, z;
        vxyz(F.d, v1, x, y, z);
        i64 best = -1;
        for (int c = 0; c < ST.adj_n[sl]; c++) {
            int k = ST.adj_l[sl][c];
            i64 nx = x + DIR[k][0], ny = y + DIR[k][1], nz = z + DIR[k][2];
            if (!inb(F.d, nx, ny, nz)) continue;
            i64 n = vidx(F.d, nx, ny, nz);
            if (fa(v1, n) != fa(v2, n)) {
                if (best < 0 || fa(best, n)) best = n;
            }
        }
        return best;
    }

    // get_smallest_vertex main_3d.cu:2115-2194: lowest f-vertex completing a tet with the face
    // (v1, v2, v3) that lies below exactly two of them, -1 if none.
    DMTZ_HD i64 smallest_vertex(i64 v1, i64 v2, i64 v3) const {
        int s2 = slot_between(F.d, v1, v2), s3 = slot_between(F.d, v1, v3);
        if (s2 < 0 || s3 < 0) return -1;
        int lt = ST.tri_of[s2][s3];
        if (lt < 0) return -1;
        i64 x, y, z;
        vxyz(F.d, v1, x, y, z);
        i64 best = -1;
        for (int c = 0; c < 2; c++) {
            int k = ST.tri_k[lt][c];
            i64 nx = x + DIR[k][0], ny = y + DIR[k][1], nz = z + DIR[k][2];
            if (!inb(F.d, nx, ny, nz)) continue;
            i64 v4 = vidx(F.d, nx, ny, nz);
            bool c1 = fa(v1, v4), c2 = fa(v2, v4), c3 = fa(v3, v4);
            if ((c1 && c2 && !c3) || (c1 && !c2 && c3) || (!c1 && c2 && c3)) {
                if (best < 0 || fa(best, v4)) best = v4;
            }
        }
        return best;
    }

    // fix_minimum main_3d.cu:2554-2586
    DMTZ_HD void fix_minimum(i64 id) {
        i64 p = F.pv(id);
        if (p != -1) {
            i64 a, b;
            edge_verts(F.d, F.c, p, a, b);
            tgt(a != id ? a : b);
        } else {
            tgt(id);
        }
    }

    // fix_saddle main_3d.cu:2588-3037
    DMTZ_HD void fix_saddle(i64 id) {
        const i64 NV = F.c.N, NF = F.c.F, NE = F.c.E;
        i64 p = F.pe(id);
        if (p != -1) {
            if (p < NV && p >= 0 && F.pv(p) == id) {  // :2602-2641
                i64 v1, v2;
                edge_verts(F.d, F.c, id, v1, v2);
                if (v2 != p) v1 = p;
                else {
                    v2 = v1;
                    v1 = p;
                }
                i64 dv = G.pv(v1);
                if (dv == -1 || dv < 0 || dv >= NE) return;
                i64 a, b;
                edge_verts(F.d, F.c, dv, a, b);
                i64 false_v2 = (v1 == a) ? b : a;
                if (ga(v2, false_v2)) tgt(v2);
            } else if (p < NF && p >= 0 && F.pf(p) == id) {  // :2643-2827
                i64 v1, v2, v3;
                edge_verts(F.d, F.c, id, v1, v2);
                tri_third(p, v1, v2, v3);
                i64 other1 = !ga(v1, v3) ? v3 : v1;
                i64 other1_1 = !ga(v2, v3) ? v3 : v2;
                bool v1e = ga(v1, v3), v2e = ga(v2, v3);
                bool o1 = fa(v1, v3), o2 = fa(v2, v3);
                if (v1e && v2e) {
                    tgt(v3);
                } else if (o1 && o2) {
                    tgt(v3);
                } else {
                    i64 q = G.pf(p);
                    if (q != -1 && G.pe(q) == p) {  // :2719-2799
                        i64 a1, a2;
                        edge_verts(F.d, F.c, q, a1, a2);
                        i64 false_v = (a1 != v1 && a1 != v2) ? a1 : a2;
                        double diff = delta(false_v);
                        i64 b3;
                        tri_third(p, a1, a2, b3);
                        bool ov1 = fa(a2, b3), ov2 = fa(a1, b3);
                        i64 pq = F.pe(q);
                        if (ov1 && ov2 && pq != -1 && F.pf(pq) == q) {
                            i64 c3;
                            tri_third(pq, a1, a2, c3);
                            tgt(c3);
                        } else if (fabs_d(diff) > 0) {
                            tgt(false_v);
                        }
                    } else {
                        if (o1 && !v1e) tgt(other1);
                        else tgt(other1_1);
                    }
                }
            }
        } else {  // false negative :2830-3036
            i64 q = G.pe(id);
            if (q != -1 && q < NV && q >= 0 && G.pv(q) == id) {
                tgt(q);
            } else if (q != -1 && q < NF && q >= 0 && G.pf(q) == id) {
                i64 v1, v2, v3;
                edge_verts(F.d, F.c, id, v1, v2);
                tri_third(q, v1, v2, v3);
                bool v1e = ga(v2, v3), v2e = ga(v1, v3);  // :2886-2887
                bool or1 = fa(v2, v3), or2 = fa(v1, v3);
                if (v1e == v2e) {
                    if (v1e != or1) tgt(v2);
                    if (v2e != or2) tgt(v1);
                } else {
                    i64 t = F.pe(id);
                    if (t != -1) {
                        if (t >= 0 && t < NV && F.pv(t) == id) tgt(v1 == t ? v2 : v1);
                    } else {
                        i64 m = lowest_between_f(v1, v2);
                        if (m != -1) tgt(m);
                        else {
                            tgt(v1);
                            tgt(v2);
                        }
                    }
                }
            }
        }
    }

    // fix_2saddle main_3d.cu:3039-3983
    DMTZ_HD void fix_2saddle(i64 id) {
        const i64 NV = F.c.N, NE = F.c.E, NF = F.c.F, NT = F.c.T;
        (void)NF;
        i64 p = F.pf(id);
        if (p != -1) {
            if (p < NE && p >= 0 && F.pe(p) == id) {  // :3057-3223
                i64 v1, v2, v3;
                edge_verts(F.d, F.c, p, v1, v2);
                tri_third(id, v1, v2, v3);
                bool v1e = ga(v2, v3), v2e = ga(v1, v3);
                bool o1 = fa(v2, v3), o2 = fa(v1, v3);
                if (v1e && v2e) {
                    i64 fid = F.pe(p);
                    if (fid == -1 || F.pf(fid) != p) return;
                    i64 c3;
                    tri_third(fid, v1, v2, c3);
                    tgt(c3);
                } else if (o1 != o2 && v1e != v2e) {
                    i64 fid = G.pe(p);
                    if (fid == -1 || G.pf(fid) != p) {
                        tgt(v3);
                    } else if (fid < F.c.F && fid >= 0) {
                        i64 w3;
                        tri_third(fid, v1, v2, w3);
                        bool e1 = ga(v2, w3), e2 = ga(v1, w3);
                        bool oe1 = fa(v2, w3), oe2 = fa(v1, w3);
                        if (e1 && e2 && ((!oe1 && oe2) || (oe1 && !oe2))) {
                            if (!oe1) tgt(v2);
                            else tgt(v1);
                        } else {
                            tgt(v3);
                        }
                    }
                } else if (!ga(v1, v3) || !ga(v2, v3)) {
                    tgt(v3);
                }
            } else if (p < NT && p >= 0 && F.pt(p) == id) {  // :3226-3477
                i64 v1, v2, v3;
                tri_verts(F.d, F.c, id, v1, v2, v3);
                i64 q1, q2, q3, q4;
                tet_verts(F.d, p, q1, q2, q3, q4);
                i64 v4 = tet_other(q1, q2, q3, q4, v1, v2, v3);
                bool or1 = fa(v3, v4), or2 = fa(v2, v4), or3 = fa(v1, v4);
                if (!or1 || !or2 || !or3) {
                    i64 pff = G.pt(p);
                    if (pff != -1) {
                        i64 w1, w2, w3;
                        tri_verts(F.d, F.c, pff, w1, w2, w3);
                        bool g1 = ga(w3, v4), g2 = ga(w2, v4), g3 = ga(w1, v4);
                        bool h1 = fa(w3, v4), h2 = fa(w2, v4), h3 = fa(w1, v4);
                        if ((!g1 || !g2 || !g3) && (!h1 || !h2 || !h3)) {
                            i64 idd = F.pf(pff);
                            if (idd >= 0 && idd < NE && F.pe(idd) == pff) {  // :3288
                                i64 e1, e2;
                                edge_verts(F.d, F.c, idd, e1, e2);
                                i64 vv3 = w3;
                                if (w1 != e1 && e2 != w1) vv3 = w1;
                                else if (w2 != e1 && w2 != e2) vv3 = w2;
                                tgt(vv3);
                            } else if (idd >= 0 && idd < NT && F.pt(idd) == pff) {  // :3302
                                i64 t1, t2, t3, t4;
                                tet_verts(F.d, idd, t1, t2, t3, t4);
                                // :3307-3309 assign the face's vertex in place of the tet's
                                if (t1 != w1 && t1 != w2 && t1 != w3) t4 = w1;
                                else if (t2 != w1 && t2 != w2 && t2 != w3) t4 = w2;
                                else if (t3 != w1 && t3 != w2 && t3 != w3) t4 = w3;
                                tgt(t1);
                                tgt(t2);
                                tgt(t3);
                                tgt(t4);
                            }
                        } else {
                            tgt(w1);
                            tgt(w2);
                            tgt(w3);
                        }
                    } else {
                        tgt(v4);
                    }
                } else {
                    tgt(v4);
                }
            }
        } else {  // false negative :3480-3981
            i64 q = G.pf(id);
            if (q != -1 && q < NE && G.pe(q) == id) {  // :3488-3610
                i64 v1, v2;
                edge_verts(F.d, F.c, q, v1, v2);
                i64 s1 = !ga(v1, v2) ? v2 : v1;
                i64 s2 = !ga(v1, v2) ? v1 : v2;
                i64 tf = F.pe(q);
                if (tf != -1 && F.pf(tf) == q) {
                    i64 a, b, v3;
                    tri_verts(F.d, F.c, tf, a, b, v3);
                    if (s1 != a && s2 != a) v3 = a;
                    if (s1 != b && s2 != b) v3 = b;
                    tgt(v3);
                } else if (tf >= 0 && tf < NV && F.pv(tf) == q) {
                    tgt(v1);
                    tgt(v2);
                }
            } else if (q != -1 && q < NT && G.pt(q) == id) {  // :3613-3980
                i64 v1, v2, v3;
                tri_verts(F.d, F.c, id, v1, v2, v3);
                i64 q1, q2, q3, q4;
                tet_verts(F.d, q, q1, q2, q3, q4);
                i64 v4 = tet_other(q1, q2, q3, q4, v1, v2, v3);
                bool f1 = ga(v3, v4), f2 = ga(v2, v4), f3 = ga(v1, v4);
                bool o1 = fa(v3, v4), o2 = fa(v2, v4), o3 = fa(v1, v4);
                if (f1 && f2 && f3) {
                    // inner else-branches of :3714-3927 are unreachable under f1 && f2 && f3
                    if (f1 != o1) tgt(v3);
                    else if (f2 != o2) tgt(v2);
                    else if (f3 != o3) tgt(v1);
                } else {
                    i64 tc = F.pt(q);
                    if (tc >= 0 && tc < NT) {  // :3937 compares a face id against the tet count
                        tgt(v1);
                        tgt(v2);
                        tgt(v3);
                    } else {
                        i64 m = smallest_vertex(v1, v2, v3);
                        if (m != -1) tgt(m);
                        else {
                            tgt(v1);
                            tgt(v2);
                            tgt(v3);
                        }
                    }
                }
            }
        }
    }

    // fix_maximum main_3d.cu:3985-4303
    DMTZ_HD void fix_maximum(i64 id) {
        const i64 NE = F.c.E, NT = F.c.T;
        i64 p = F.pt(id);
        if (p != -1) {
            i64 v1, v2, v3;
            tri_verts(F.d, F.c, p, v1, v2, v3);
            i64 q1, q2, q3, q4;
            tet_verts(F.d, id, q1, q2, q3, q4);
            i64 v4 = tet_other(q1, q2, q3, q4, v1, v2, v3);
            bool o1 = fa(v1, v4), o2 = fa(v2, v4), o3 = fa(v3, v4);
            if (o1 && o2 && o3) {
                tgt(v4);
            } else {
                i64 fc = G.pf(p);
                if (fc >= 0) {
                    if (fc < NE && G.pe(fc) == p) {
                        i64 a, b;
                        edge_verts(F.d, F.c, fc, a, b);
                        tgt(ga(a, b) ? a : b);
                    } else if (fc < NT && G.pt(fc) == p) {
                        tgt(v4);
                    }
                }
            }
        } else {
            i64 face1 = G.pt(id);
            i64 v1, v2, v3;
            tri_verts(F.d, F.c, face1, v1, v2, v3);
            i64 q1, q2, q3, q4;
            tet_verts(F.d, id, q1, q2, q3, q4);
            i64 v4 = tet_other(q1, q2, q3, q4, v1, v2, v3);
            bool f1 = ga(v1, v4), f2 = ga(v2, v4), f3 = ga(v3, v4);
            bool o1 = fa(v1, v4), o2 = fa(v2, v4), o3 = fa(v3, v4);
            if (f1 && f2 && f3) {
                if (!o1) tgt(v1);
                if (!o2) tgt(v2);
                if (!o3) tgt(v3);
                if (o1 && o2 && o3) {
                    i64 it = F.pf(face1);
                    if (it != -1) {
                        i64 t1, t2, t3, t4;
                        tet_verts(F.d, it, t1, t2, t3, t4);
                        tgt(tet_other(t1, t2, t3, t4, v1, v2, v3));
                    }
                }
            } else {
                i64 pid = F.pf(face1);
                if (pid != -1) {
                    if (pid >= 0 && pid < NE && F.pe(pid) == face1) {
                        i64 a, b;
                        edge_verts(F.d, F.c, pid, a, b);
                        tgt(a);
                        tgt(b);
                    } else {
                        i64 t1, t2, t3, t4;
                        tet_verts(F.d, pid, t1, t2, t3, t4);
                        tgt(t1);
                        tgt(t2);
                        tgt(t3);
                        tgt(t4);
                    }
                } else {
                    if (!o1) tgt(v1);
                    if (!o2) tgt(v2);
                    if (!o3) tgt(v3);
                }
            }
        }
    }

    // fix_as_vpath main_3d.cu:5135-5244, live branch; t_idx = the divergent tet
    DMTZ_HD void fix_as_vpath(i64 t_idx) {
        i64 te = F.pt(t_idx);
        i64 fe = G.pt(t_idx);
        if (fe == -1 || te == -1) return;
        i64 pfe = F.pf(fe);
        if (pfe >= 0 && pfe < F.c.T && F.pt(pfe) == fe) {
            i64 v1, v2, v3;
            tri_verts(F.d, F.c, fe, v1, v2, v3);
            i64 q1, q2, q3, q4;
            tet_verts(F.d, pfe, q1, q2, q3, q4);
            tgt(tet_other(q1, q2, q3, q4, v1, v2, v3));
        } else {
            i64 v1, v2, v3;
            tri_verts(F.d, F.c, te, v1, v2, v3);
            i64 a, b, tv3;
            tri_verts(F.d, F.c, fe, a, b, tv3);
            if (a != v1 && a != v2 && a != v3) tv3 = a;
            else if (b != v1 && b != v2 && b != v3) tv3 = b;
            tgt(tv3);
        }
    }

    // fix_ds_vpath main_3d.cu:5246-5295, live branch; v_idx = the recorded vertex
    DMTZ_HD void fix_ds_vpath(i64 v_idx) {
        if (v_idx < 0) return;
        i64 e = F.pv(v_idx);
        if (e == -1) return;
        i64 a, b;
        edge_verts(F.d, F.c, e, a, b);
        tgt(v_idx == a ? b : a);
    }

    // fix_ds_wall main_3d.cu:5297-5413; e = the first edge of the wall whose f and g partners
    // differ. The partner of an edge is a face id or, for an edge paired with its vertex, that
    // vertex id; the range checks below compare both kinds against the face count, as the
    // reference does. The first branch's third-vertex pick is two independent ifs (:5331-5334),
    // the last branch's an else-if.
    DMTZ_HD void fix_ds_wall(i64 e) {
        const i64 NF = F.c.F;
        const i64 tf = F.pe(e), ff = G.pe(e);
        if (tf == ff) return;
        i64 v1, v2;
        edge_verts(F.d, F.c, e, v1, v2);
        if (tf >= 0 && ff >= 0 && tf < NF && F.pf(tf) == e && ff < NF && G.pf(ff) == e) {
            i64 a, b, v3;
            tri_verts(F.d, F.c, tf, a, b, v3);
            if (a != v1 && a != v2) v3 = a;
            if (b != v1 && b != v2) v3 = b;
            if (fabs_d(delta(v3)) > 1e-16 * range) {
                tgt(v3);
            } else {
                tgt(v1);
                tgt(v2);
            }
        } else if (tf == -1 || tf >= NF || (tf >= 0 && F.pf(tf) != e)) {
            tgt(v1);
            tgt(v2);
        } else if (ff == -1 || ff >= NF || (ff >= 0 && G.pf(ff) != e)) {
            i64 a, b, v3;
            tri_verts(F.d, F.c, tf, a, b, v3);
            if (a != v1 && a != v2) v3 = a;
            else if (b != v1 && b != v2) v3 = b;
            tgt(v3);
        }
    }

    // Face partners of the three edges of face t in f's pairing, for the wall walk
    // (get_false_ds_wall main_3d.cu:5064-5133): edges (v1,v2), (v1,v3), (v2,v3) of the
    // descending vertex triple. mismatch = the first edge whose f partner exists and differs
    // from its g partner (-1 if none before the stop); next[k] = the face paired with edge k in
    // f when that face is not t and pairs back to edge k, else -1. Edges after the mismatch are
    // not examined (next = -1).
    DMTZ_HD void wall_step(i64 t, i64& mismatch, i64 next[3]) const {
        i64 v1, v2, v3;
        tri_verts(F.d, F.c, t, v1, v2, v3);
        auto edge = [&](i64 a, i64 b) {
            const int sl = slot_between(F.d, a, b);
            return sl < 0 ? (i64)-1 : F.edge_of(a, sl);
        };
        const i64 es[3] = {edge(v1, v2), edge(v1, v3), edge(v2, v3)};
        mismatch = -1;
        for (int k = 0; k < 3; k++) {
            next[k] = -1;
            if (mismatch >= 0) continue;
            const i64 j = es[k];
            const i64 pj = F.pe(j);
            if (pj == -1) continue;
            if (pj != G.pe(j)) {
                mismatch = j;
                continue;
            }
            if (pj != t && pj >= 0 && pj < F.c.F && F.pf(pj) == j) next[k] = pj;
        }
    }

    // The wall walk from face `face` (get_false_ds_wall main_3d.cu:5064-5133 with a stack and
    // visited set of its own): depth first, a face is marked when popped, the faces of
    // wall_step are pushed in edge order so the last is popped first; found = the first edge
    // met whose f and g partners differ, -1 if the wall has none. The visited set is open
    // addressing over `slots` entries (a power of two; face id + 1, 0 = empty) in set[], the
    // stack is stack[0 .. slots + 2). Returns true (overflow, found undefined) when the walk
    // would mark more than slots / 2 faces; each marked face pushes at most 3, so the stack
    // then holds at most slots + 1. f's gradient has no closed V-path, so every walk ends.
    // Face ids must be below 2^32 - 1.
    DMTZ_HD bool wall_walk(i64 face, uint32_t* set, uint32_t* stack, uint32_t slots, i64& found) const {
        for (uint32_t k = 0; k < slots; k++) set[k] = 0;
        const uint32_t mask = slots - 1, limit = slots / 2, shift = 32 - (uint32_t)__builtin_ctz(slots);
        uint32_t sp = 0, used = 0;
        stack[sp++] = (uint32_t)face;
        found = -1;
        bool overflow = false;
        while (sp > 0 && found < 0 && !overflow) {
            const uint32_t tri = stack[--sp];
            uint32_t h = (uint32_t)(tri * 2654435761u) >> shift;  // multiplicative hash, top bits
            bool seen = false;
            for (uint32_t probe = 0; probe < slots; probe++) {  // the set is at most half full
                const uint32_t key = set[h];
                if (key == tri + 1u) {
                    seen = true;
                    break;
                }
                if (key == 0) break;
                h = (h + 1) & mask;
            }
            if (seen) continue;
            if (used == limit) {
                overflow = true;
                continue;
            }
            set[h] = tri + 1u;
            used++;
            i64 mismatch, next[3];
            wall_step((i64)tri, mismatch, next);
            if (mismatch >= 0) {
                found = mismatch;
                continue;
            }
            for (int k = 0; k < 3; k++) {
                if (next[k] < 0) continue;
                if (sp >= slots + 2) {
                    overflow = true;
                    break;
                }
                stack[sp++] = (uint32_t)next[k];
            }
        }
        return overflow;
    }

    // ---------------------------------------------------------------------------------------
    // Separatrix checks
    // ---------------------------------------------------------------------------------------
    // Cofacet tets of face t in the reference's search order (main_3d.cu:4380-4426):
    // x = first, y = second or -1.
    DMTZ_HD void cofacets(i64 t, i64& x, i64& y) const {
        i64 v1, v2, v3;
        tri_verts(F.d, F.c, t, v1, v2, v3);
        x = -1;
        y = -1;
        int s2 = slot_between(F.d, v1, v2), s3 = slot_between(F.d, v1, v3);
        if (s2 < 0 || s3 < 0) return;
        int lt = ST.tri_of[s2][s3];
        if (lt < 0) return;
        i64 px, py, pz;
        vxyz(F.d, v1, px, py, pz);
        int n = 0;
        for (int c = 0; c < 2; c++) {
            int k = ST.tri_k[lt][c];
            i64 nx = px + DIR[k][0], ny = py + DIR[k][1], nz = pz + DIR[k][2];
            if (!inb(F.d, nx, ny, nz)) continue;
            i64 q = tet_id4(F.d, F.c, v1, v2, v3, vidx(F.d, nx, ny, nz));
            if (n == 0) x = q;
            else y = q;
            n++;
        }
    }

    // Walks are written as one loop with a completion flag and no early exit from nested
    // loops: the shape the device compiler handles reliably for these long divergent loops.

    // is_false main_3d.cu:4437-4497: walk f's ascending path from tet id; on the first tet
    // whose f and g partner faces differ, found = true and divergent = that tet. Capped at cap
    // steps (overflow = true).
    DMTZ_HD void as_walk(i64 id, i64 cap, bool& found, i64& divergent, bool& overflow) const {
        i64 cur = id;
        bool done = id < 0;
        found = false;
        divergent = -1;
        i64 steps = 0;
        while (!done) {
            i64 t = F.pt(cur);
            i64 dt = t == -1 ? -1 : G.pt(cur);
            i64 pft = (t == -1 || t != dt) ? -1 : F.pf(t);
            bool chain = pft != -1 && F.pt(pft) == t;
            i64 t1 = -1, t2 = -1;
            if (chain) cofacets(t, t1, t2);
            i64 nxt = (t1 == cur) ? t2 : t1;
            if (t != -1 && t != dt) {
                found = true;
                divergent = cur;
            }
            done = t == -1 || t != dt || !chain || nxt == cur || nxt == -1;
            cur = nxt;
            if (!done && ++steps > cap) {
                overflow = true;
                done = true;
            }
        }
    }

    // The comparison steps of as_walk from tet id, taken on f alone: every tet cur_i of f's
    // ascending path whose f partner face t_i is not -1, in walk order, up to the step where the
    // chain ends. as_walk(id) finds the first step with G.pt(cur_i) != t_i (divergent = cur_i,
    // found for the start tet id), or nothing when all steps agree: f is fixed during the
    // correction, so these steps are computed once per run. out(i, cur_i, t_i) receives each
    // step; returns the number of steps (overflow = true past cap steps).
    template <class Out>
    DMTZ_HD i64 as_fpath(i64 id, i64 cap, bool& overflow, Out out) const {
        i64 cur = id, n = 0;
        bool done = id < 0;
        while (!done) {
            done = as_fpath_step(cur, n, out);
            if (!done && n > cap) {
                overflow = true;
                done = true;
            }
        }
        return n;
    }

    // One comparison step of as_fpath from tet cur (n steps so far): records (n, cur, t), moves
    // cur to the next tet and returns true when the path ends there. as_walk's chain test
    // (F.pf(t) names a tet whose F.pt is t again) holds whenever t = F.pt(cur) != -1: a tet and
    // the triangle it pairs with are paired in one step of the gradient, each with the other
    // (main_3d.cu:2310-2311), so F.pf(t) = cur. The step therefore reads one pairing instead of
    // three.
    template <class Out>
    DMTZ_HD bool as_fpath_step(i64& cur, i64& n, Out out) const {
        const i64 t = F.pt(cur);
        i64 nxt = -1;
        if (t != -1) {
            out(n, cur, t);
            n++;
            i64 t1 = -1, t2 = -1;
            cofacets(t, t1, t2);
            nxt = (t1 == cur) ? t2 : t1;
        }
        const bool done = t == -1 || nxt == cur || nxt == -1;
        cur = nxt;
        return done;
    }

    // get_paired_id main_3d.cu:4499-4537 on field which (0 = f, 1 = g): the tet reached by
    // following the ascending pairing from tet id.
    DMTZ_HD i64 paired_max(i64 id, int which, i64 cap, bool& overflow) const {
        const FieldGrad<T>& X = which == 0 ? F : G;
        i64 cur = id;
        bool done = id < 0;
        i64 steps = 0;
        while (!done) {
            i64 t = X.pt(cur);
            i64 pft = t == -1 ? -1 : X.pf(t);
            bool chain = pft != -1 && X.pt(pft) == t;
            i64 t1 = -1, t2 = -1;
            if (chain) cofacets(t, t1, t2);
            i64 nxt = (t1 == cur) ? t2 : t1;
            done = !chain || nxt == cur || nxt == -1;
            if (chain) cur = nxt;  // leaving through the boundary yields -1, as the reference
            if (!done && ++steps > cap) {
                overflow = true;
                done = true;
            }
        }
        return cur;
    }

    // get_false_ds_vpath main_3d.cu:4837-4959 for one f-critical edge. The reference's
    // dec_current_id (:4849) never advances, so each walk compares only its start vertex:
    // v1's f and g pairings, then v2's; the terminal f minimum is not compared.
    DMTZ_HD bool ds_geometry(i64 e, i64& rec) const {
        i64 v1, v2;
        edge_verts(F.d, F.c, e, v1, v2);
        i64 c1 = F.pv(v1), c2 = F.pv(v2);
        bool b1 = c1 != -1 && c1 != G.pv(v1);
        bool b2 = c2 != -1 && c2 != G.pv(v2);
        rec = b1 ? v1 : v2;
        return b1 || b2;
    }

    // get_orginal_saddle_min_pairs main_3d.cu:4613-4687: the f minimum reached from start.
    DMTZ_HD i64 f_min_label(i64 start, i64 cap, bool& overflow) const {
        i64 cur = start;
        bool done = start < 0;
        i64 steps = 0;
        while (!done) {
            i64 ce = F.pv(cur);
            bool chain = ce != -1 && F.pe(ce) != -1;
            i64 a = -1, b = -1;
            if (chain) edge_verts(F.d, F.c, ce, a, b);
            done = !chain;
            if (chain) cur = (a == cur) ? b : a;
            if (!done && ++steps > cap) {
                overflow = true;
                done = true;
            }
        }
        return cur;
    }

    // get_false_ds_vpath_connectivity main_3d.cu:4961-5062, one endpoint: walk g's descending
    // path, keep the first vertex whose f and g pairings differ (trouble stays if already set),
    // return the g minimum reached.
    DMTZ_HD i64 g_min_walk(i64 start, i64& trouble, i64 cap, bool& overflow) const {
        i64 cur = start;
        bool done = start < 0;
        i64 steps = 0;
        while (!done) {
            i64 dce = G.pv(cur);
            i64 fce = dce == -1 ? -1 : F.pv(cur);
            if (dce != -1 && fce != dce && trouble == -1) trouble = cur;
            bool chain = dce != -1 && G.pe(dce) != -1;
            i64 a = -1, b = -1;
            if (chain) edge_verts(F.d, F.c, dce, a, b);
            done = !chain;
            if (chain) cur = (a == cur) ? b : a;
            if (!done && ++steps > cap) {
                overflow = true;
                done = true;
            }
        }
        return cur;
    }
};

}  // namespace dmtz
