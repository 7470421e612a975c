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
    uint32_t pe;        // edges paired with the centre (V) or a triangle (E0 / E1)
    uint32_t blw[14];   // blw[i] bit j: n_j below n_i

    DMTZ_HD int e0_of(int a) const { return (int)((e0 >> (4 * a)) & 15u); }
    DMTZ_HD int e1_of(int a) const { return (int)((e1 >> (4 * a)) & 15u); }
    DMTZ_HD uint32_t crit_edges() const { return L & ~pe; }
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

// The lowest element of S (a subset of the star slots) as a one-hot mask, 0 if S is empty:
// the element of S with nothing of S below it. Candidates are restricted to the compile-time
// list of the slot's link.
DMTZ_HD uint32_t lowest_bit(uint32_t S, const uint32_t* blw, const uint8_t* cand, int ncand) {
    uint32_t m = 0;
    for (int c = 0; c < ncand; c++) {
        const int k = cand[c];
        m |= ((S >> k) & (uint32_t)((blw[k] & S) == 0u) & 1u) << k;
    }
    return m;
}

// Fresh copies of the slot masks for one pass: bit tests inside the pass are not shared with
// other passes, so they do not stay live between them.
struct PassMasks {
    uint32_t b[14];
    uint32_t L;
    DMTZ_HD PassMasks(const uint32_t* blw, uint32_t L0) {
#pragma unroll
        for (int i = 0; i < 14; i++) {
            b[i] = blw[i];
            DMTZ_OPAQUE(b[i]);
        }
        L = L0;
        DMTZ_OPAQUE(L);
    }
};

DMTZ_HD uint32_t bit1(uint32_t x, int k) { return (x >> k) & 1u; }
DMTZ_HD uint32_t bit1(uint64_t x, int k) { return (uint32_t)(x >> k) & 1u; }

}  // namespace detail

// Evaluate the five passes on the lower star of the centre of val[15] (val[14] = centre).
// Every per-cell condition is computed as 0/1 integer arithmetic on the slot masks, with the
// selections resolved at compile time: "below both n_i and n_j" is blw[i] & blw[j], "strictly
// between them" is blw[i] ^ blw[j] (the order is total, so whatever is below the lower of the
// two is below the higher), and the triangles paired in E0 / E1 are set from the slot each edge
// chose, through the compile-time triangle index of (edge slot, chosen slot).
template <class T>
DMTZ_HD void lower_star(const T* val, LowerStar& r) {
    using detail::bit1;
    uint32_t blw[14];
#pragma unroll
    for (int i = 0; i < 14; i++) blw[i] = 0;
    detail::PairLoop<0, 1>::run(val, blw);
    uint32_t L = 0;
    detail::CentreLoop<0>::run(val, L);
#pragma unroll
    for (int i = 0; i < 14; i++) DMTZ_OPAQUE(blw[i]);
    DMTZ_OPAQUE(L);
    r.L = L;
#pragma unroll
    for (int i = 0; i < 14; i++) r.blw[i] = blw[i];

    // V pass: steepest descent (main_3d.cu:1953-1975)
    uint32_t smask = 0;
#pragma unroll
    for (int k = 0; k < 14; k++) smask |= (bit1(L, k) & (uint32_t)((blw[k] & L) == 0u)) << k;
    r.s = smask ? (int)__builtin_ctz(smask) : -1;
    const uint32_t Ls = L & ~smask;  // edges of the lower star not taken by V

    // E0 (main_3d.cu:2057-2071): edge {v, n_a} with a != s takes its lowest link vertex below n_a.
    // The triangle of (a, chosen slot) is paired with the edge.
    uint64_t e0 = 0xFFull << 56, tE0 = 0;  // bits above the 14 nibbles stay set
    uint32_t pe0 = 0;  // edges paired in E0
    {
    const detail::PassMasks P(blw, L);
#pragma unroll
    for (int a = 0; a < 14; a++) {
        const uint32_t Ca = P.L & ST.adj[a] & P.b[a];
        uint32_t m = detail::lowest_bit(Ca, P.b, ST.adj_l[a], ST.adj_n[a]);
        m &= 0u - bit1(Ls, a);
        uint32_t nib = 15;
#pragma unroll
        for (int c = 0; c < 6; c++) {
            if (c < ST.adj_n[a]) {
                const int k = ST.adj_l[a][c];
                const uint32_t sel = bit1(m, k);
                nib ^= (15u ^ (uint32_t)k) & (0u - sel);
                tE0 |= (uint64_t)sel << ST.adj_t[a][c];
            }
        }
        e0 |= (uint64_t)nib << (4 * a);
        pe0 |= (uint32_t)(m != 0u) << a;
        DMTZ_OPAQUE(e0);
        DMTZ_OPAQUE(tE0);
        DMTZ_OPAQUE(pe0);
    }
    }
    r.e0 = e0;

    // Triangles: membership, orientation and T0 (main_3d.cu:2269-2279). T0 can only reach a
    // vertex below both other vertices of the triangle, so the triangle is the top face of the
    // tet it chooses and no other face can choose that tet: the tet is claimed exactly when its
    // top face pairs with it, and the claims are collected here.
    uint64_t tin = 0, thi = 0, tT0 = 0, tD0 = 0;
    uint32_t qT0 = 0;
    {
    const detail::PassMasks P(blw, L);
#pragma unroll
    for (int t = 0; t < 36; t++) {
        const int i = ST.tri_i[t], j = ST.tri_j[t];
        const int k0 = ST.tri_k[t][0], k1 = ST.tri_k[t][1];
        const uint32_t in = bit1(P.L, i) & bit1(P.L, j);
        const uint32_t BB = P.L & P.b[i] & P.b[j];
        const uint32_t c0 = bit1(BB, k0), c1 = bit1(BB, k1);
        const uint32_t pick1 = c1 & ((c0 ^ 1u) | bit1(P.b[k0], k1));
        const uint32_t pick0 = c0 & (pick1 ^ 1u);
        const uint32_t fr = in & (bit1(tE0, t) ^ 1u);
        const uint32_t p1 = fr & pick1, p0 = fr & pick0;
        qT0 |= (p0 << ST.tri_q[t][0]) | (p1 << ST.tri_q[t][1]);
        tin |= (uint64_t)in << t;
        thi |= (uint64_t)bit1(P.b[i], j) << t;  // n_j below n_i
        tT0 |= (uint64_t)(p0 | p1) << t;
        tD0 |= (uint64_t)p1 << t;
        DMTZ_OPAQUE(qT0);
        DMTZ_OPAQUE(tin);
        DMTZ_OPAQUE(thi);
        DMTZ_OPAQUE(tT0);
        DMTZ_OPAQUE(tD0);
    }
    }

    // E1 (main_3d.cu:2073-2096): a free edge {v, n_a} takes the lowest link vertex between
    // n_a and v, if that triangle is still free after E0 and T0.
    const uint64_t before = tE0 | tT0;
    const uint32_t freeE = Ls & ~pe0;
    uint64_t e1 = 0xFFull << 56, tE1 = 0;
    uint32_t pe1 = 0;  // edges paired in E1
    {
    const detail::PassMasks P(blw, L);
#pragma unroll
    for (int a = 0; a < 14; a++) {
        const uint32_t Ma = P.L & ST.adj[a] & ~P.b[a];
        uint32_t m = detail::lowest_bit(Ma, P.b, ST.adj_l[a], ST.adj_n[a]);
        m &= 0u - bit1(freeE, a);
        uint32_t nib = 15;
#pragma unroll
        for (int c = 0; c < 6; c++) {
            if (c < ST.adj_n[a]) {
                const int k = ST.adj_l[a][c];
                const int t = ST.adj_t[a][c];
                const uint32_t ok = bit1(m, k) & (bit1(before, t) ^ 1u);
                nib ^= (15u ^ (uint32_t)k) & (0u - ok);
                tE1 |= (uint64_t)ok << t;
                pe1 |= ok << a;
            }
        }
        e1 |= (uint64_t)nib << (4 * a);
        DMTZ_OPAQUE(e1);
        DMTZ_OPAQUE(tE1);
        DMTZ_OPAQUE(pe1);
    }
    }
    r.e1 = e1;

    // Tets of the lower star.
    uint32_t qin = 0;
#pragma unroll
    for (int q = 0; q < 24; q++)
        qin |= (bit1(L, ST.tet_s[q][0]) & bit1(L, ST.tet_s[q][1]) & bit1(L, ST.tet_s[q][2])) << q;

    // Triangles: T1 (main_3d.cu:2282-2299). T1 reaches the vertex strictly between the
    // triangle's other two, so the triangle is the face of the chosen tet without its middle
    // vertex, the only face that can claim it in T1: the claims are collected here.
    const uint64_t taken = before | tE1;
    uint64_t tT1 = 0, tD1 = 0;
    uint32_t qT1 = 0;
    {
    const detail::PassMasks P(blw, L);
#pragma unroll
    for (int t = 0; t < 36; t++) {
        const int i = ST.tri_i[t], j = ST.tri_j[t];
        const int k0 = ST.tri_k[t][0], k1 = ST.tri_k[t][1];
        const int q0 = ST.tri_q[t][0], q1 = ST.tri_q[t][1];
        const uint32_t BT = P.L & (P.b[i] ^ P.b[j]);
        const uint32_t c0 = bit1(BT, k0), c1 = bit1(BT, k1);
        const uint32_t pick1 = c1 & ((c0 ^ 1u) | bit1(P.b[k0], k1));
        const uint32_t pick0 = c0 & (pick1 ^ 1u);
        const uint32_t fr = bit1(tin, t) & (bit1(taken, t) ^ 1u);
        const uint32_t p1 = fr & pick1 & (bit1(qT0, q1) ^ 1u);
        const uint32_t p0 = fr & pick0 & (bit1(qT0, q0) ^ 1u);
        qT1 |= (p0 << q0) | (p1 << q1);
        tT1 |= (uint64_t)(p0 | p1) << t;
        tD1 |= (uint64_t)p1 << t;
        DMTZ_OPAQUE(qT1);
        DMTZ_OPAQUE(tT1);
        DMTZ_OPAQUE(tD1);
    }
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
    r.pe = smask | pe0 | pe1;
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

namespace detail {

// Bit i of the low 21 bits of x moved to bit 3i (Morton-code bit spreading).
DMTZ_HD uint64_t spread3(uint64_t x) {
    x &= 0x1fffffull;
    x = (x | x << 32) & 0x1f00000000ffffull;
    x = (x | x << 16) & 0x1f0000ff0000ffull;
    x = (x | x << 8) & 0x100f00f00f00f00full;
    x = (x | x << 4) & 0x10c30c30c30c30c3ull;
    x = (x | x << 2) & 0x1249249249249249ull;
    return x;
}

// Bit i of the low 32 bits of x moved to bit 2i.
DMTZ_HD uint64_t spread2(uint64_t x) {
    x &= 0xffffffffull;
    x = (x | x << 16) & 0x0000ffff0000ffffull;
    x = (x | x << 8) & 0x00ff00ff00ff00ffull;
    x = (x | x << 4) & 0x0f0f0f0f0f0f0f0full;
    x = (x | x << 2) & 0x3333333333333333ull;
    x = (x | x << 1) & 0x5555555555555555ull;
    return x;
}

}  // namespace detail

// The code's fields are built as bit planes over all edges / triangles / tets at once; the pass
// results are mutually exclusive per cell (each cell is paired at most once), so each plane is an
// OR of masked pass masks.
DMTZ_HD void encode_star(const LowerStar& r, StarCode& out) {
    // Edges: the E0 and E1 nibbles are 15 except where that pass paired the edge (never both),
    // so the partner nibble is their AND; the steepest slot (15 in both) gets 14.
    uint64_t w0 = r.e0 & r.e1 & ((1ull << 56) - 1ull);
    if (r.s >= 0) w0 &= ~(1ull << (4 * r.s));
    w0 |= (uint64_t)(r.s < 0 ? 15 : r.s) << 56;

    // Triangles: 0/1 = edge to tri_i / tri_j, 2/3 = tet via tri_k[t][0] / [1], 7 = unpaired.
    const uint64_t T36 = (1ull << 36) - 1ull;
    const uint64_t un = ~(r.tE0 | r.tT0 | r.tE1 | r.tT1) & T36;
    const uint64_t b2 = un;
    const uint64_t b1 = (r.tT0 | r.tT1 | un) & T36;
    const uint64_t b0 = ((r.tE0 & ~r.thi) | (r.tE1 & r.thi) | (r.tT0 & r.tD0) | (r.tT1 & r.tD1) | un) & T36;
    const uint64_t w2 = detail::spread3(b0) | (detail::spread3(b1) << 1) | (detail::spread3(b2) << 2);
    const uint64_t w3 =
        detail::spread3(b0 >> 21) | (detail::spread3(b1 >> 21) << 1) | (detail::spread3(b2 >> 21) << 2);

    // Tets: T0 leaves out the lowest slot of (tet_s[q][0], [1], [2]), T1 the middle one; 3 =
    // unpaired. ji / ki / kj: n_j below n_i, n_k below n_i, n_k below n_j.
    uint32_t JI = 0, KI = 0, KJ = 0;
#pragma unroll
    for (int q = 0; q < 24; q++) {
        const int i = ST.tet_s[q][0], j = ST.tet_s[q][1], k = ST.tet_s[q][2];
        JI |= ((r.blw[i] >> j) & 1u) << q;
        KI |= ((r.blw[i] >> k) & 1u) << q;
        KJ |= ((r.blw[j] >> k) & 1u) << q;
        DMTZ_OPAQUE(JI);
        DMTZ_OPAQUE(KI);
        DMTZ_OPAQUE(KJ);
    }
    const uint32_t low1 = KI & KJ, low0 = JI & ~KJ;
    const uint32_t midi = JI ^ KI, midj = ~JI ^ KJ;
    const uint32_t mid0 = ~midi & midj, mid1 = ~midi & ~midj;
    const uint32_t fr = ~(r.qT0 | r.qT1) & 0xffffffu;
    const uint32_t c0 = ((r.qT0 & low0) | (r.qT1 & mid0) | fr) & 0xffffffu;
    const uint32_t c1 = ((r.qT0 & low1) | (r.qT1 & mid1) | fr) & 0xffffffu;
    const uint64_t w1 = (uint64_t)r.L | ((detail::spread2(c0) | (detail::spread2(c1) << 1)) << 14);

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
