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
static int grad_impl(const T* fld, i64 W, i64 H, i64 D, i64* pv, i64* pe, i64* pf, i64* pt) {
    Dims d{W, H, D};
    Counts c = make_counts(d);
    FieldGrad<T> g(fld, d, c);
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
    Repair<T, VecSink> R(f, g, d, c, counter.data(), eps, range, edit_type, VecSink{&tg});
    for (int k = 0; k < 4; k++) counts4[k] = 0;
    std::vector<std::pair<int, i64>> recs;
    FieldGrad<T> Fg(f, d, c), Gg(g, d, c);
    for (i64 v = 0; v < c.N; v++) {
        T val[15];
        LowerStar a, b;
        gather_star(f, d, v, val);
        lower_star(val, a);
        gather_star(g, d, v, val);
        lower_star(val, b);
        uint64_t lo;
        uint32_t hi;
        vertex_mismatch(Fg, Gg, v, a, b, lo, hi);
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
