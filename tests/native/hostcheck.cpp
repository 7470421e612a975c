// Host build of the product's device-agnostic headers (lower_star.h, gradient.h, repair.h) for
// unit tests on machines without a GPU. Test harness only: it compiles the same header code the
// HIP kernels use with g++, so the lower-star classifier, the pairing lookups and the repair
// rules can be compared with the oracle on the CPU. It is never loaded by the product path.
#include <cstdint>
#include <cstring>
#include <vector>
#include "repair.h"
#include "mismatch.h"

using namespace dmtz;

template <class T>
static std::vector<StarCode> all_codes(const T* fld, const Dims& d, const Counts& c) {
    std::vector<StarCode> out(c.N);
    for (i64 v = 0; v < c.N; v++) {
        LowerStar ls;
        star_of(fld, d, v, ls);
        encode_star(ls, out[v]);
    }
    return out;
}

template <class T>
static int grad_impl(const T* fld, i64 W, i64 H, i64 D, i64* pv, i64* pe, i64* pf, i64* pt) {
    Dims d{W, H, D};
    Counts c = make_counts(d);
    std::vector<StarCode> cd = all_codes(fld, d, c);
    FieldGrad<T> g(fld, cd.data(), d, c);
    for (i64 v = 0; v < c.N; v++) pv[v] = g.pv(v);
    for (i64 e = 0; e < c.E; e++) pe[e] = g.pe(e);
    for (i64 t = 0; t < c.F; t++) pf[t] = g.pf(t);
    for (i64 q = 0; q < c.T; q++) pt[q] = g.pt(q);
    return 0;
}

extern "C" int hc_gradient_f32(const float* f, i64 W, i64 H, i64 D, i64* pv, i64* pe, i64* pf, i64* pt) {
    return grad_impl<float>(f, W, H, D, pv, pe, pf, pt);
}
extern "C" int hc_gradient_f64(const double* f, i64 W, i64 H, i64 D, i64* pv, i64* pe, i64* pf, i64* pt) {
    return grad_impl<double>(f, W, H, D, pv, pe, pf, pt);
}

// Critical masks per vertex (75 local cells): lo 51 bits, hi 24 bits.
template <class T>
static int crit_impl(const T* fld, i64 W, i64 H, i64 D, uint64_t* lo, uint32_t* hi) {
    Dims d{W, H, D};
    Counts c = make_counts(d);
    for (i64 v = 0; v < c.N; v++) {
        T val[15];
        gather_star(fld, d, v, val);
        LowerStar ls;
        lower_star(val, ls);
        crit_mask(ls, lo[v], hi[v]);
    }
    return 0;
}
extern "C" int hc_crit_f32(const float* f, i64 W, i64 H, i64 D, uint64_t* lo, uint32_t* hi) {
    return crit_impl<float>(f, W, H, D, lo, hi);
}

struct VecSink {
    std::vector<uint8_t>* t;
    void operator()(i64 v) const { (*t)[v] = 1; }
};

template <class T>
static int targets_impl(const T* f, const T* g, i64 W, i64 H, i64 D, double eps, double range, int preserve_saddles,
                        int edit_type, uint8_t* targets, i64* counts4) {
    Dims d{W, H, D};
    Counts c = make_counts(d);
    std::vector<uint8_t> tg(c.N, 0);
    std::vector<uint8_t> counter(c.N, 1);
    std::vector<StarCode> cf = all_codes(f, d, c), cg = all_codes(g, d, c);
    Repair<T, VecSink> R(f, cf.data(), g, cg.data(), d, c, counter.data(), eps, range, edit_type, VecSink{&tg});
    for (int k = 0; k < 4; k++) counts4[k] = 0;
    std::vector<std::pair<int, i64>> recs;
    FieldGrad<T> Fg(f, cf.data(), d, c);
    for (i64 v = 0; v < c.N; v++) {
        uint64_t flo, lo;
        uint32_t fhi, hi;
        local_crit(Fg, v, flo, fhi);
        owner_mismatch(cg[v], flo, fhi, lo, hi);
        for (int lc = 0; lc < LC_COUNT; lc++) {
            if (!lc_bit(lo, hi, lc)) continue;
            int ty = lc_type(lc);
            counts4[ty]++;
            recs.push_back({ty, local_cell_id(d, c, v, lc)});
        }
    }
    for (auto& r : recs) {
        if (r.first == 0) R.fix_minimum(r.second);
        else if (r.first == 1 && preserve_saddles) R.fix_saddle(r.second);
        else if (r.first == 2 && preserve_saddles) R.fix_2saddle(r.second);
        else if (r.first == 3) R.fix_maximum(r.second);
    }
    std::memcpy(targets, tg.data(), c.N);
    return 0;
}
extern "C" int hc_targets_f32(const float* f, const float* g, i64 W, i64 H, i64 D, double eps, double range,
                              int ps, int et, uint8_t* targets, i64* counts4) {
    return targets_impl<float>(f, g, W, H, D, eps, range, ps, et, targets, counts4);
}
extern "C" int hc_targets_f64(const double* f, const double* g, i64 W, i64 H, i64 D, double eps, double range,
                              int ps, int et, uint8_t* targets, i64* counts4) {
    return targets_impl<double>(f, g, W, H, D, eps, range, ps, et, targets, counts4);
}

// Separatrix mismatch lists (geometry or connectivity mode) on (f, g) as they stand, with the
// same per-cell logic as the device kernels: descending over f-critical edges, ascending over
// f-critical faces. Outputs (id, recorded) pairs; returns -1 when a walk hit the cap.
struct NoSinkH {
    void operator()(i64) const {}
};
template <class T>
static int paths_impl(const T* f, const T* g, i64 W, i64 H, i64 D, int geometry, i64* counts2, i64* ds_id,
                      i64* ds_rec, i64* as_id, i64* as_rec) {
    Dims d{W, H, D};
    Counts c = make_counts(d);
    std::vector<StarCode> cf = all_codes(f, d, c), cg = all_codes(g, d, c);
    Repair<T, NoSinkH> R(f, cf.data(), g, cg.data(), d, c, nullptr, 0, 0, 0, NoSinkH{});
    Repair<T, NoSinkH> RF(f, cf.data(), f, cf.data(), d, c, nullptr, 0, 0, 0, NoSinkH{});
    bool ovf = false;
    i64 nds = 0, nas = 0;
    for (i64 v = 0; v < c.N; v++) {
        uint64_t lo, ilo;
        uint32_t hi, ihi;
        code_masks(cf[v], lo, hi, ilo, ihi);
        for (int lc = 1; lc < LC_TET; lc++) {
            if (!lc_bit(lo, hi, lc)) continue;
            i64 id = local_cell_id(d, c, v, lc);
            if (lc < LC_TRI) {
                i64 rec = -1;
                if (geometry) {
                    if (R.ds_geometry(id, rec)) { ds_id[nds] = id; ds_rec[nds++] = rec; }
                } else {
                    i64 a, b;
                    edge_verts(d, c, id, a, b);
                    i64 la = RF.f_min_label(a, c.N, ovf), lb = RF.f_min_label(b, c.N, ovf);
                    i64 trouble = -1;
                    i64 end = R.g_min_walk(a, trouble, c.N, ovf);
                    if (end != la) { ds_id[nds] = id; ds_rec[nds++] = trouble; }
                    else {
                        end = R.g_min_walk(b, trouble, c.N, ovf);
                        if (end != lb) { ds_id[nds] = id; ds_rec[nds++] = trouble; }
                    }
                }
            } else {
                i64 q[2];
                R.cofacets(id, q[0], q[1]);
                for (int k = 0; k < 2; k++) {
                    if (q[k] < 0) continue;
                    if (!geometry) {
                        i64 lab = RF.paired_max(q[k], 0, c.N, ovf);
                        if (R.paired_max(q[k], 1, c.N, ovf) == lab) continue;
                    }
                    i64 dv;
                    bool found;
                    R.as_walk(q[k], c.N, found, dv, ovf);
                    if (found) { as_id[nas] = q[k]; as_rec[nas++] = dv; break; }
                }
            }
        }
    }
    counts2[0] = nds;
    counts2[1] = nas;
    return ovf ? -1 : 0;
}
extern "C" int hc_paths_f32(const float* f, const float* g, i64 W, i64 H, i64 D, int geometry, i64* counts2,
                            i64* ds_id, i64* ds_rec, i64* as_id, i64* as_rec) {
    return paths_impl<float>(f, g, W, H, D, geometry, counts2, ds_id, ds_rec, as_id, as_rec);
}

// Edit set of the separatrix repairs over given (tet, divergent tet) and (edge, vertex) lists.
template <class T>
static int path_targets_impl(const T* f, const T* g, i64 W, i64 H, i64 D, double eps, double range, int edit_type,
                             const i64* as_rec, i64 nas, const i64* ds_rec, i64 nds, uint8_t* targets) {
    Dims d{W, H, D};
    Counts c = make_counts(d);
    std::vector<uint8_t> tg(c.N, 0);
    std::vector<uint8_t> counter(c.N, 1);
    std::vector<StarCode> cf = all_codes(f, d, c), cg = all_codes(g, d, c);
    Repair<T, VecSink> R(f, cf.data(), g, cg.data(), d, c, counter.data(), eps, range, edit_type, VecSink{&tg});
    for (i64 k = 0; k < nas; k++) R.fix_as_vpath(as_rec[k]);
    for (i64 k = 0; k < nds; k++) R.fix_ds_vpath(ds_rec[k]);
    std::memcpy(targets, tg.data(), c.N);
    return 0;
}
extern "C" int hc_path_targets_f32(const float* f, const float* g, i64 W, i64 H, i64 D, double eps, double range,
                                   int et, const i64* as_rec, i64 nas, const i64* ds_rec, i64 nds, uint8_t* targets) {
    return path_targets_impl<float>(f, g, W, H, D, eps, range, et, as_rec, nas, ds_rec, nds, targets);
}

// Ascending geometry check of every (f-critical face, cofacet) branch two ways: Repair::as_walk
// directly, and as the first step of Repair::as_fpath (computed on f alone) whose g partner face
// differs. out_walk / out_steps receive (found tet, divergent tet) per branch; returns the
// number of branches.
struct StepVec {
    std::vector<std::pair<i64, i64>>* v;
    void operator()(i64, i64 cur, i64 t) const { v->push_back({cur, t}); }
};
template <class T>
static i64 as_two_ways_impl(const T* f, const T* g, i64 W, i64 H, i64 D, i64* out_walk, i64* out_steps) {
    Dims d{W, H, D};
    Counts c = make_counts(d);
    std::vector<StarCode> cf = all_codes(f, d, c), cg = all_codes(g, d, c);
    Repair<T, NoSinkH> R(f, cf.data(), g, cg.data(), d, c, nullptr, 0, 0, 0, NoSinkH{});
    Repair<T, NoSinkH> RF(f, cf.data(), f, cf.data(), d, c, nullptr, 0, 0, 0, NoSinkH{});
    FieldGrad<T> G(g, cg.data(), d, c);
    i64 nb = 0;
    for (i64 v = 0; v < c.N; v++) {
        uint64_t lo, ilo;
        uint32_t hi, ihi;
        code_masks(cf[v], lo, hi, ilo, ihi);
        for (int lc = LC_TRI; lc < LC_TET; lc++) {
            if (!lc_bit(lo & ilo, hi & ihi, lc)) continue;
            i64 t = local_cell_id(d, c, v, lc), q[2];
            R.cofacets(t, q[0], q[1]);
            for (int k = 0; k < 2; k++) {
                bool ovf = false, found = false;
                i64 dv = -1;
                R.as_walk(q[k], c.N, found, dv, ovf);
                out_walk[2 * nb] = found ? q[k] : -1;
                out_walk[2 * nb + 1] = dv;
                std::vector<std::pair<i64, i64>> st;
                RF.as_fpath(q[k], c.N, ovf, StepVec{&st});
                i64 fq = -1, fdv = -1;
                for (auto& p : st)
                    if (G.pt(p.first) != p.second) {
                        fq = st[0].first;
                        fdv = p.first;
                        break;
                    }
                out_steps[2 * nb] = fq;
                out_steps[2 * nb + 1] = fdv;
                nb++;
            }
        }
    }
    return nb;
}
extern "C" i64 hc_as_two_ways_f32(const float* f, const float* g, i64 W, i64 H, i64 D, i64* ow, i64* os) {
    return as_two_ways_impl<float>(f, g, W, H, D, ow, os);
}

// FastDiv / divmod against native 64-bit division: divisors 1..4096 and large 32-bit ones, numerators
// at the edges (0, d-1, d, multiples, 2^63-1, 2^64-1) and pseudo-random. Returns the number of
// disagreements.
extern "C" i64 hc_divmod_check(i64 samples) {
    i64 bad = 0;
    uint64_t s = 0x9E3779B97F4A7C15ull;
    auto rnd = [&]() {
        s ^= s << 13;
        s ^= s >> 7;
        s ^= s << 17;
        return s;
    };
    std::vector<uint64_t> ds;
    for (uint64_t d = 1; d <= 4096; d++) ds.push_back(d);
    for (int k = 0; k < 2000; k++) ds.push_back(1 + (rnd() >> (33 + rnd() % 31)));
    ds.push_back(511ull * 512);
    ds.push_back(1023ull * 1023);
    ds.push_back((1ull << 31) - 1);
    ds.push_back((1ull << 32) - 1);
    for (uint64_t d : ds) {
        FastDiv fd((uint32_t)d);
        uint64_t edge[] = {0, 1, d - 1, d, d + 1, 2 * d - 1, 2 * d, 1000 * d + 7, (1ull << 63) - 1,
                           ~0ull, ~0ull - d, (1ull << 34) + 12345};
        for (uint64_t n : edge) bad += fd.div(n) != n / d;
        for (i64 k = 0; k < samples; k++) {
            uint64_t n = rnd() >> (rnd() % 64);
            bad += fd.div(n) != n / d;
            i64 sn = (i64)(rnd() >> 1) * ((rnd() & 1) ? -1 : 1) >> (rnd() % 62), q, r;
            divmod(sn, fd, q, r);
            bad += (q != sn / (i64)d) + (r != sn % (i64)d);
        }
    }
    return bad;
}

// The ID decoders of stencil.h for a list of ids (negative and past-the-end ones included):
// out receives 3 vertices per triangle id, then 4 per tet id.
extern "C" void hc_decode(i64 W, i64 H, i64 D, const i64* tris, i64 nt, const i64* tets, i64 nq, i64* out) {
    Dims d{W, H, D};
    Counts c = make_counts(d);
    for (i64 k = 0; k < nt; k++) tri_verts(d, c, tris[k], out[3 * k], out[3 * k + 1], out[3 * k + 2]);
    i64* o = out + 3 * nt;
    for (i64 k = 0; k < nq; k++) tet_verts(d, tets[k], o[4 * k], o[4 * k + 1], o[4 * k + 2], o[4 * k + 3]);
}

// f-critical masks of every vertex three ways: by lookups at each cell's owner (local_crit),
// scattered (each owner spreads its critical cells to the cells' other vertices) and gathered
// from the neighbours' owned critical cells (crit_gather). Returns the number of vertices whose
// scattered or gathered masks differ from local_crit's.
template <class T>
static i64 critf_two_ways_impl(const T* f, i64 W, i64 H, i64 D) {
    Dims d{W, H, D};
    Counts c = make_counts(d);
    std::vector<StarCode> cf = all_codes(f, d, c);
    FieldGrad<T> F(f, cf.data(), d, c);
    std::vector<uint64_t> slo(c.N);
    std::vector<uint32_t> shi(c.N);
    for (i64 v = 0; v < c.N; v++) {
        uint64_t lo, ilo;
        uint32_t hi, ihi;
        code_masks(cf[v], lo, hi, ilo, ihi);
        slo[v] = lo & ilo;
        shi[v] = hi & ihi;
    }
    const std::vector<uint64_t> olo = slo;
    const std::vector<uint32_t> ohi = shi;
    for (i64 v = 0; v < c.N; v++)
        crit_scatter_spread(d, cf[v], v, [&](i64 w, int lc) {
            if (lc < 51) slo[w] |= 1ull << lc;
            else shi[w] |= 1u << (lc - 51);
        });
    i64 bad = 0;
    for (i64 v = 0; v < c.N; v++) {
        uint64_t lo, glo;
        uint32_t hi, ghi;
        local_crit(F, v, lo, hi);
        crit_gather(d, v, [&](i64 u, uint64_t& a, uint32_t& b) { a = olo[u]; b = ohi[u]; }, glo, ghi);
        bad += (lo != slo[v]) || (hi != shi[v]) || (lo != glo) || (hi != ghi);
    }
    return bad;
}
extern "C" i64 hc_critf_two_ways_f32(const float* f, i64 W, i64 H, i64 D) { return critf_two_ways_impl<float>(f, W, H, D); }
extern "C" i64 hc_critf_two_ways_f64(const double* f, i64 W, i64 H, i64 D) { return critf_two_ways_impl<double>(f, W, H, D); }

// Walls from every f-critical face with Repair::wall_walk (the code the wall kernel runs), with
// the kernel's tiers of visited-set size; then the fix_ds_wall edit set over the list. Returns
// the number of walls (face/edge receive them in face order), -1 if a walk outgrows every tier.
template <class T>
static i64 walls_impl(const T* f, const T* g, i64 W, i64 H, i64 D, double eps, double range, int edit_type,
                      i64* face, i64* edge, uint8_t* targets) {
    Dims d{W, H, D};
    Counts c = make_counts(d);
    std::vector<StarCode> cf = all_codes(f, d, c), cg = all_codes(g, d, c);
    std::vector<uint8_t> tg(c.N, 0);
    std::vector<uint8_t> counter(c.N, 1);
    Repair<T, VecSink> R(f, cf.data(), g, cg.data(), d, c, counter.data(), eps, range, edit_type, VecSink{&tg});
    const FieldGrad<T> F(f, cf.data(), d, c);
    i64 n = 0;
    for (i64 t = 0; t < c.F; t++) {
        if (F.pf(t) != -1) continue;
        i64 found = -1;
        bool overflow = true;
        for (uint32_t slots : {256u, 8192u, 1u << 18}) {
            std::vector<uint32_t> buf(2 * (size_t)slots + 2);
            overflow = R.wall_walk(t, buf.data(), buf.data() + slots, slots, found);
            if (!overflow) break;
        }
        if (overflow) return -1;
        if (found >= 0) {
            face[n] = t;
            edge[n] = found;
            n++;
        }
    }
    for (i64 k = 0; k < n; k++) R.fix_ds_wall(edge[k]);
    std::memcpy(targets, tg.data(), c.N);
    return n;
}
extern "C" i64 hc_walls_f32(const float* f, const float* g, i64 W, i64 H, i64 D, double eps, double range, int et,
                            i64* face, i64* edge, uint8_t* targets) {
    return walls_impl<float>(f, g, W, H, D, eps, range, et, face, edge, targets);
}
