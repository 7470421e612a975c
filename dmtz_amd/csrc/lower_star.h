// Per-vertex lower-star form of the reference's five-pass discrete gradient.
//
// The reference computes the gradient with five dense passes over all simplices
// (main_3d.cu:1930-2113 and 2197-2316, launched in the order of main_3d.cu:6864-6874):
//   V : vertex v pairs with edge {v, s}, s = lowest lower neighbour          (:1953-1984)
//   E0: edge {a,b} pairs with triangle {a,b,c}, c = lowest common neighbour
//       below both, if the triangle is free                                  (:2057-2071, 2103)
//   T0: triangle {a,b,c} pairs with tet {a,b,c,d}, d = lowest common
//       neighbour below all three, if the tet is free                         (:2269-2279, 2307)
//   E1: a still-free edge pairs with {a,n,b}, n = lowest common neighbour
//       strictly between its endpoints, if the triangle is free              (:2073-2096)
//   T1: a still-free triangle pairs with the tet through the lowest d below
//       exactly two of its vertices, if the tet is free                      (:2282-2299)
// In every pass the added vertex is below the cell's highest vertex, so every pair lies in the
// lower star of one vertex, and each pass has a unique claimant per target cell. The gradient
// restricted to LowerStar(v) is therefore a function of the 15 values of v's closed star, and
// this header evaluates it per vertex with bitmask logic over the 14 star slots.
//
// Order: u < w  <=>  val[u] < val[w] or (val[u] == val[w] and index u < index w)
// (main_3d.cu:1043-1049). Inside one star the index order of two slots is the fixed rank TB[].
// Absent (off-grid) neighbours carry +inf and so are never below anything; fields are finite.
#pragma once
#include "stencil.h"

namespace dmtz {

// n_i below n_j, for compile-time slots i != j (slot 14 = centre).
template <class T, int I, int J>
DMTZ_HD bool below(const T* val) {
    if constexpr (TB[I] < TB[J]) return val[I] <= val[J];
    else return val[I] < val[J];
}
template <class T>
DMTZ_HD bool below_rt(const T* val, int i, int j) {
    return TB[i] < TB[j] ? (val[i] <= val[j]) : (val[i] < val[j]);
}

// Result of the five passes restricted to the lower star of one vertex.
struct LowerStar {
    uint32_t L;         // slots below the centre
    int s;              // V: steepest slot, -1 if the centre is a minimum
    uint64_t e0;        // E0: 4 bits per slot, third slot of the edge's triangle, 15 = none
    uint64_t e1;        // E1: 4 bits per slot, middle slot of the edge's triangle, 15 = none
    uint64_t tin;       // triangles (36 bits) in the lower star
    uint64_t thi;       // orientation: bit t set when n_{tri_i} is above n_{tri_j}
    uint64_t tE0, tT0, tE1, tT1;  // triangle paired in E0 / T0 / E1 / T1
    uint64_t tD0, tD1;  // tet chosen in T0 / T1 is via tri_k[t][1] (else tri_k[t][0])
    uint32_t qin;       // tets (24 bits) in the lower star
    uint32_t qT0, qT1;  // tet claimed in T0 / T1
    uint32_t blw[14];   // blw[i] bit j: n_j below n_i

    DMTZ_HD int e0_of(int a) const { return (int)((e0 >> (4 * a)) & 15u); }
    DMTZ_HD int e1_of(int a) const { return (int)((e1 >> (4 * a)) & 15u); }
    DMTZ_HD uint32_t crit_edges() const {
        uint32_t m = 0;
        for (int a = 0; a < 14; a++) {
            bool c = ((L >> a) & 1u) && a != s && e0_of(a) == 15 && e1_of(a) == 15;
            m |= (uint32_t)c << a;
        }
        return m;
    }
    DMTZ_HD uint64_t crit_tris() const { return tin & ~(tE0 | tT0 | tE1 | tT1); }
    DMTZ_HD uint32_t crit_tets() const { return qin & ~(qT0 | qT1); }
};

namespace detail {

template <int I, int J>
struct PairLoop {
    template <class T>
    DMTZ_HD static void run(const T* val, uint32_t* blw) {
        bool c = below<T, I, J>(val);  // n_I below n_J
        blw[J] |= c ? (1u << I) : 0u;
        blw[I] |= c ? 0u : (1u << J);
        if constexpr (J + 1 < 14) PairLoop<I, J + 1>::run(val, blw);
        else if constexpr (I + 2 < 14) PairLoop<I + 1, I + 2>::run(val, blw);
    }
};

template <int I>
struct CentreLoop {
    template <class T>
    DMTZ_HD static void run(const T* val, uint32_t& L) {
        L |= below<T, I, 14>(val) ? (1u << I) : 0u;
        if constexpr (I + 1 < 14) CentreLoop<I + 1>::run(val, L);
    }
};

// Lowest element of S (S a subset of the star slots), or 15 if S is empty. Candidates are
// restricted to the compile-time list of the slot's link.
DMTZ_HD int lowest_in(uint32_t S, const uint32_t* blw, const uint8_t* cand, int ncand) {
    int m = 15;
    for (int c = 0; c < ncand; c++) {
        int k = cand[c];
        bool is = ((S >> k) & 1u) && (blw[k] & S) == 0u;
        m = is ? k : m;
    }
    return m;
}

}  // namespace detail

// Evaluate the five passes on the lower star of the centre of val[15] (val[14] = centre).
template <class T>
DMTZ_HD void lower_star(const T* val, LowerStar& r) {
    uint32_t blw[14];
#pragma unroll
    for (int i = 0; i < 14; i++) blw[i] = 0;
    detail::PairLoop<0, 1>::run(val, blw);
    uint32_t L = 0;
    detail::CentreLoop<0>::run(val, L);
    r.L = L;
#pragma unroll
    for (int i = 0; i < 14; i++) r.blw[i] = blw[i];

    // V pass: steepest descent (main_3d.cu:1953-1975)
    int s = -1;
#pragma unroll
    for (int k = 0; k < 14; k++) {
        bool is = ((L >> k) & 1u) && (blw[k] & L) == 0u;
        s = is ? k : s;
    }
    r.s = s;

    // E0 (main_3d.cu:2057-2071): edge {v, n_a} with a != s takes its lowest link vertex below n_a.
    uint64_t e0 = ~0ull;
#pragma unroll
    for (int a = 0; a < 14; a++) {
        uint32_t Ca = L & ST.adj[a] & blw[a];
        int c = detail::lowest_in(Ca, blw, ST.adj_l[a], ST.adj_n[a]);
        bool ok = ((L >> a) & 1u) && a != s && c != 15;
        uint64_t nib = ok ? (uint64_t)c : 15ull;
        e0 &= ~(15ull << (4 * a));
        e0 |= nib << (4 * a);
    }
    r.e0 = e0;

    // Triangles: orientation, E0 status, T0 (main_3d.cu:2269-2279). T0 can only reach a vertex
    // below both other vertices of the triangle, so the triangle is the top face of the tet it
    // chooses and no other face can choose that tet: the tet is claimed exactly when its top face
    // pairs with it, and the claims are collected here.
    uint64_t tin = 0, thi = 0, tE0 = 0, tT0 = 0, tD0 = 0;
    uint32_t qT0 = 0;
#pragma unroll
    for (int t = 0; t < 36; t++) {
        const int i = ST.tri_i[t], j = ST.tri_j[t];
        const int k0 = ST.tri_k[t][0], k1 = ST.tri_k[t][1];
        bool in = ((L >> i) & 1u) && ((L >> j) & 1u);
        bool hi = (blw[i] >> j) & 1u;  // n_j below n_i: a = i, b = j
        int a = hi ? i : j, b = hi ? j : i;
        int e0a = (int)((e0 >> (4 * a)) & 15u);
        bool pE0 = in && e0a == b;
        uint32_t blwb = hi ? blw[j] : blw[i];
        bool c0 = ((L >> k0) & 1u) && ((blwb >> k0) & 1u);
        bool c1 = ((L >> k1) & 1u) && ((blwb >> k1) & 1u);
        bool pick1 = c1 && (!c0 || ((blw[k0] >> k1) & 1u));
        bool pT0 = in && !pE0 && (c0 || c1);
        qT0 |= (uint32_t)pT0 << (pick1 ? ST.tri_q[t][1] : ST.tri_q[t][0]);
        tin |= (uint64_t)in << t;
        thi |= (uint64_t)hi << t;
        tE0 |= (uint64_t)pE0 << t;
        tT0 |= (uint64_t)pT0 << t;
        tD0 |= (uint64_t)(pT0 && pick1) << t;
    }

    // E1 (main_3d.cu:2073-2096): a free edge {v, n_a} takes the lowest link vertex between
    // n_a and v, if that triangle is still free after E0 and T0.
    uint64_t e1 = ~0ull;
#pragma unroll
    for (int a = 0; a < 14; a++) {
        bool freeE = ((L >> a) & 1u) && a != s && (int)((e0 >> (4 * a)) & 15u) == 15;
        uint32_t Ma = L & ST.adj[a] & ~blw[a];
        int m = 15;
        bool taken = false;
#pragma unroll
        for (int c = 0; c < 6; c++) {
            if (c < ST.adj_n[a]) {
                const int k = ST.adj_l[a][c];
                const int t = ST.adj_t[a][c];
                bool is = ((Ma >> k) & 1u) && (blw[k] & Ma) == 0u;
                m = is ? k : m;
                taken = is ? (((tE0 | tT0) >> t) & 1ull) != 0 : taken;
            }
        }
        bool ok = freeE && m != 15 && !taken;
        uint64_t nib = ok ? (uint64_t)m : 15ull;
        e1 &= ~(15ull << (4 * a));
        e1 |= nib << (4 * a);
    }
    r.e1 = e1;

    // Tets of the lower star.
    uint32_t qin = 0;
#pragma unroll
    for (int q = 0; q < 24; q++) {
        const int i = ST.tet_s[q][0], j = ST.tet_s[q][1], k = ST.tet_s[q][2];
        qin |= (uint32_t)(((L >> i) & 1u) && ((L >> j) & 1u) && ((L >> k) & 1u)) << q;
    }

    // Triangles: E1 status and T1 (main_3d.cu:2282-2299). T1 reaches the vertex strictly between
    // the triangle's other two, so the triangle is the face of the chosen tet without its middle
    // vertex, the only face that can claim it in T1: the claims are collected here.
    uint64_t tE1 = 0, tT1 = 0, tD1 = 0;
    uint32_t qT1 = 0;
#pragma unroll
    for (int t = 0; t < 36; t++) {
        const int i = ST.tri_i[t], j = ST.tri_j[t];
        const int k0 = ST.tri_k[t][0], k1 = ST.tri_k[t][1];
        bool in = (tin >> t) & 1ull;
        bool hi = (thi >> t) & 1ull;
        int a = hi ? i : j, b = hi ? j : i;
        bool before = (((tE0 | tT0) >> t) & 1ull) != 0;
        int e1b = (int)((e1 >> (4 * b)) & 15u);
        bool pE1 = in && !before && e1b == a;
        uint32_t blwa = hi ? blw[i] : blw[j];
        uint32_t blwb = hi ? blw[j] : blw[i];
        // strictly between n_b and n_a
        bool c0 = ((L >> k0) & 1u) && ((blwa >> k0) & 1u) && !((blwb >> k0) & 1u);
        bool c1 = ((L >> k1) & 1u) && ((blwa >> k1) & 1u) && !((blwb >> k1) & 1u);
        bool pick1 = c1 && (!c0 || ((blw[k0] >> k1) & 1u));
        const int q0 = ST.tri_q[t][0], q1 = ST.tri_q[t][1];
        bool tetFree = pick1 ? !((qT0 >> q1) & 1u) : !((qT0 >> q0) & 1u);
        bool pT1 = in && !before && !pE1 && (c0 || c1) && tetFree;
        qT1 |= (uint32_t)pT1 << (pick1 ? q1 : q0);
        tE1 |= (uint64_t)pE1 << t;
        tT1 |= (uint64_t)pT1 << t;
        tD1 |= (uint64_t)(pT1 && pick1) << t;
    }


    r.tin = tin;
    r.thi = thi;
    r.tE0 = tE0;
    r.tT0 = tT0;
    r.tE1 = tE1;
    r.tT1 = tT1;
    r.tD0 = tD0;
    r.tD1 = tD1;
    r.qin = qin;
    r.qT0 = qT0;
    r.qT1 = qT1;
}

// Local cell numbering of the 75 cells whose highest vertex is the centre:
//   0 = the vertex, 1..14 = edge to slot a (1 + a), 15..50 = star triangle t (15 + t),
//   51..74 = star tet q (51 + q).
enum : int { LC_VERT = 0, LC_EDGE = 1, LC_TRI = 15, LC_TET = 51, LC_COUNT = 75 };

// Critical-cell mask of a lower star in the local numbering, split in two words:
// lo = vertex + 14 edges + 36 triangles (bits 0..50), hi = 24 tets.
DMTZ_HD void crit_mask(const LowerStar& r, uint64_t& lo, uint32_t& hi) {
    lo = (uint64_t)(r.L == 0u) | ((uint64_t)r.crit_edges() << 1) | (r.crit_tris() << 15);
    hi = r.crit_tets();
}

}  // namespace dmtz

namespace dmtz {

// Compact form of a lower star's pairing, 32 bytes per vertex:
//   w[0] bits 0..55  partner nibble of the edge to slot a (a = 0..13): 14 = the centre vertex,
//                    0..13 = the triangle completed by that slot, 15 = critical / not in the star;
//        bits 56..59 steepest slot s (15 = none)
//   w[1] bits 0..13  L (slots below the centre); bits 14..61: 2 bits per star tet q: the slot of
//                    (tet_s[q][0], [1], [2]) its partner face leaves out, 3 = critical / absent
//   w[2] 21 triangles x 3 bits, w[3] 15 triangles x 3 bits: 0 = edge to tri_i, 1 = edge to
//                    tri_j, 2 = tet through tri_k[t][0], 3 = tet through tri_k[t][1],
//                    7 = critical / absent
struct StarCode {
    uint64_t w[4];
};

DMTZ_HD int code_edge(const StarCode& c, int a) { return (int)((c.w[0] >> (4 * a)) & 15u); }
DMTZ_HD int code_s(const StarCode& c) { return (int)((c.w[0] >> 56) & 15u); }
DMTZ_HD uint32_t code_L(const StarCode& c) { return (uint32_t)(c.w[1] & 0x3FFFu); }
DMTZ_HD int code_tet(const StarCode& c, int q) { return (int)((c.w[1] >> (14 + 2 * q)) & 3u); }
DMTZ_HD int code_tri(const StarCode& c, int t) {
    return t < 21 ? (int)((c.w[2] >> (3 * t)) & 7u) : (int)((c.w[3] >> (3 * (t - 21))) & 7u);
}

DMTZ_HD void encode_star(const LowerStar& r, StarCode& out) {
    uint64_t w0 = 0, w1 = (uint64_t)r.L, w2 = 0, w3 = 0;
#pragma unroll
    for (int a = 0; a < 14; a++) {
        int e0 = r.e0_of(a), e1 = r.e1_of(a);
        int nib = a == r.s ? 14 : (e0 != 15 ? e0 : (e1 != 15 ? e1 : 15));
        nib = ((r.L >> a) & 1u) ? nib : 15;
        w0 |= (uint64_t)nib << (4 * a);
    }
    w0 |= (uint64_t)(r.s < 0 ? 15 : r.s) << 56;
#pragma unroll
    for (int t = 0; t < 36; t++) {
        bool hi = (r.thi >> t) & 1ull;  // tri_i above tri_j
        int code = 7;
        if ((r.tE0 >> t) & 1ull) code = hi ? 0 : 1;        // edge to the higher slot
        else if ((r.tT0 >> t) & 1ull) code = 2 + (int)((r.tD0 >> t) & 1ull);
        else if ((r.tE1 >> t) & 1ull) code = hi ? 1 : 0;   // edge to the lower slot
        else if ((r.tT1 >> t) & 1ull) code = 2 + (int)((r.tD1 >> t) & 1ull);
        if (t < 21) w2 |= (uint64_t)code << (3 * t);
        else w3 |= (uint64_t)code << (3 * (t - 21));
    }
#pragma unroll
    for (int q = 0; q < 24; q++) {
        const int i = ST.tet_s[q][0], j = ST.tet_s[q][1], k = ST.tet_s[q][2];
        bool ji = (r.blw[i] >> j) & 1u, ki = (r.blw[i] >> k) & 1u, kj = (r.blw[j] >> k) & 1u;
        // T0: the face without the lowest slot; T1: the face without the middle slot
        int low = (ki && kj) ? 2 : ((ji && !kj) ? 1 : 0);
        bool midi = ji != ki, midj = (!ji) != kj;
        int mid = midi ? 0 : (midj ? 1 : 2);
        int code = 3;
        if ((r.qT0 >> q) & 1u) code = low;
        else if ((r.qT1 >> q) & 1u) code = mid;
        w1 |= (uint64_t)code << (14 + 2 * q);
    }
    out.w[0] = w0;
    out.w[1] = w1;
    out.w[2] = w2;
    out.w[3] = w3;
}

// Critical and membership masks over the 75 local cells, from a code.
DMTZ_HD void code_masks(const StarCode& c, uint64_t& crit_lo, uint32_t& crit_hi, uint64_t& in_lo, uint32_t& in_hi) {
    uint32_t L = code_L(c);
    uint64_t elo = 0, tin = 0, tcrit = 0;
#pragma unroll
    for (int a = 0; a < 14; a++) elo |= (uint64_t)(((L >> a) & 1u) && code_edge(c, a) == 15) << a;
#pragma unroll
    for (int t = 0; t < 36; t++) {
        bool in = ((L >> ST.tri_i[t]) & 1u) && ((L >> ST.tri_j[t]) & 1u);
        tin |= (uint64_t)in << t;
        tcrit |= (uint64_t)(in && code_tri(c, t) == 7) << t;
    }
    uint32_t qin = 0, qcrit = 0;
#pragma unroll
    for (int q = 0; q < 24; q++) {
        bool in = ((L >> ST.tet_s[q][0]) & 1u) && ((L >> ST.tet_s[q][1]) & 1u) && ((L >> ST.tet_s[q][2]) & 1u);
        qin |= (uint32_t)in << q;
        qcrit |= (uint32_t)(in && code_tet(c, q) == 3) << q;
    }
    crit_lo = (uint64_t)(L == 0u) | (elo << 1) | (tcrit << 15);
    crit_hi = qcrit;
    in_lo = 1ull | ((uint64_t)L << 1) | (tin << 15);
    in_hi = qin;
}

}  // namespace dmtz
