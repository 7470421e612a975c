// Freudenthal 14-neighbour stencil tables and the reference's global simplex numbering.
//
// The implicit complex is the 6-tet Freudenthal triangulation of the voxel grid with the
// stencil of main_3d.cu:44-51. Simplex IDs follow the reference exactly, because they key every
// parity output: edges main_3d.cu:917-1012 / 1760-1900, triangles main_3d.cu:246-565,
// tetrahedra main_3d.cu:572-740. The formulas here are written from the cube-local form of
// those tables: every simplex has an anchor (its componentwise-minimum corner) and a set of
// cube-local corners (bit = dx + 2 dy + 4 dz), and the ID is block offset + anchor cell index
// times the block multiplicity + a type number.
//
// Slot numbering of the star of a vertex v: slots 0..13 are the neighbours v + DIR[slot]
// (the reference's direction order), slot 14 is v itself.
#pragma once
#include <cstdint>

#if defined(__HIPCC__)
#define DMTZ_HD __host__ __device__ __forceinline__
#define DMTZ_HD_CALL __host__ __device__ __attribute__((noinline))
#else
#define DMTZ_HD inline
#define DMTZ_HD_CALL __attribute__((noinline))
#endif

// Scheduling fence for long unrolled bitmask code: keeps the compiler from hoisting the
// independent work of later iterations (which multiplies the live registers); no effect on host.
#if defined(__HIP_DEVICE_COMPILE__)
#define DMTZ_SCHED_FENCE() __builtin_amdgcn_sched_barrier(0)
#else
#define DMTZ_SCHED_FENCE() ((void)0)
#endif

// Value barrier: the compiler treats x as an unknown value from here on (no instruction is
// emitted). Used on bitmasks built from many comparisons, so that later bit tests read the mask
// instead of keeping every comparison result alive.
#if defined(__HIP_DEVICE_COMPILE__)
#define DMTZ_OPAQUE(x) asm volatile("" : "+v"(x))
#else
#define DMTZ_OPAQUE(x) ((void)0)
#endif

namespace dmtz {

typedef int64_t i64;

DMTZ_HD uint64_t mulhi64(uint64_t a, uint64_t b) { return (uint64_t)(((unsigned __int128)a * b) >> 64); }

// Unsigned division by a run-time invariant divisor d >= 1 through a precomputed multiplier
// (Granlund and Montgomery, "Division by invariant integers using multiplication", PLDI 1994,
// Figure 4.1): with l = ceil(log2 d) and m = floor(2^64 (2^l - d) / d) + 1,
// n / d = (t + ((n - t) >> min(l, 1))) >> max(l - 1, 0) where t = mulhi(m, n), for all 64-bit n.
struct FastDiv {
    uint64_t m = 1;
    uint32_t d = 1;
    uint16_t s1 = 0, s2 = 0;
    FastDiv() = default;
    explicit FastDiv(uint32_t dv) : d(dv) {
        int l = 0;
        while (l < 32 && ((uint64_t)1 << l) < dv) l++;
        const unsigned __int128 num = (unsigned __int128)(((uint64_t)1 << l) - dv) << 64;
        m = (uint64_t)(num / dv) + 1;
        s1 = (uint16_t)(l < 1 ? l : 1);
        s2 = (uint16_t)(l > 1 ? l - 1 : 0);
    }
    DMTZ_HD uint64_t div(uint64_t n) const {
        const uint64_t t = mulhi64(m, n);
        return (t + ((n - t) >> s1)) >> s2;
    }
};

// Negative numerators (ids of -1 passed through a decoder) take C's truncating division.
DMTZ_HD void divmod_neg(i64 n, i64 dv, i64& q, i64& r) {
    q = n / dv;
    r = n % dv;
}

// q = n / fd.d and r = n % fd.d with C semantics (truncation toward zero).
DMTZ_HD void divmod(i64 n, const FastDiv& fd, i64& q, i64& r) {
    if (n >= 0) {
        q = (i64)fd.div((uint64_t)n);
        r = n - q * (i64)fd.d;
    } else {
        divmod_neg(n, (i64)fd.d, q, r);
    }
}

// Grid sides plus their divisors for the ID decoders (mixed-radix digits), prepared once per
// grid on the host.
struct Dims {
    i64 W = 0, H = 0, D = 0;
    FastDiv fW, fH, fW1, fH1, fD1;
    Dims() = default;
    Dims(i64 w, i64 h, i64 dd)
        : W(w), H(h), D(dd), fW((uint32_t)w), fH((uint32_t)h), fW1((uint32_t)(w - 1)), fH1((uint32_t)(h - 1)),
          fD1((uint32_t)(dd - 1)) {}
};

// main_3d.cu:44-51
constexpr int DIR[14][3] = {{1, 0, 0},  {-1, 0, 0}, {0, 1, 0},  {0, -1, 0}, {0, 0, 1},
                            {0, 0, -1}, {-1, 1, 0}, {1, -1, 0}, {0, 1, 1},  {0, -1, -1},
                            {-1, 0, 1}, {1, 0, -1}, {-1, 1, 1}, {1, -1, -1}};

// Index-order rank of each slot's linear offset dx + dy*W + dz*W*H (slot 14 = offset 0).
// Valid for W >= 3 and H >= 2; the API rejects smaller grids. Ties between values are broken by
// global index (main_3d.cu:1043-1049), and for two vertices of one star the index order is this
// fixed rank order.
constexpr int TB[15] = {8, 6, 10, 4, 12, 2, 9, 5, 14, 0, 11, 3, 13, 1, 7};

constexpr bool dir_adjacent(int i, int j) {
    int dx = DIR[j][0] - DIR[i][0], dy = DIR[j][1] - DIR[i][1], dz = DIR[j][2] - DIR[i][2];
    for (int k = 0; k < 14; k++)
        if (DIR[k][0] == dx && DIR[k][1] == dy && DIR[k][2] == dz) return true;
    return false;
}

struct StarTables {
    uint16_t adj[14];       // adj[i] bit j: n_i and n_j are 14-neighbours
    uint8_t adj_n[14];      // |adj[i]| (4 or 6)
    uint8_t adj_l[14][6];   // the slots of adj[i]
    uint8_t tri_i[36], tri_j[36];   // star triangles (v, n_i, n_j), i < j
    uint8_t tri_k[36][2];           // the two slots completing a star tet with (i, j)
    uint8_t tri_q[36][2];           // star tet index of (i, j, tri_k[t][c])
    int8_t tri_of[14][14];          // triangle index of a slot pair, -1 if not adjacent
    uint8_t tet_s[24][3];           // star tets (v, n_i, n_j, n_k), i < j < k
    uint8_t tet_f[24][3];           // triangle index of faces (i,j), (i,k), (j,k)
    uint8_t adj_t[14][6];           // triangle index of (i, adj_l[i][c])
    int8_t slot_of[27];             // slot of offset (dx+1)*9 + (dy+1)*3 + (dz+1), -1 if none
};

constexpr StarTables make_star_tables() {
    StarTables s{};
    for (int i = 0; i < 14; i++) {
        s.adj[i] = 0;
        s.adj_n[i] = 0;
        for (int j = 0; j < 14; j++)
            if (j != i && dir_adjacent(i, j)) {
                s.adj[i] |= (uint16_t)(1u << j);
                s.adj_l[i][s.adj_n[i]++] = (uint8_t)j;
            }
    }
    for (int i = 0; i < 14; i++)
        for (int j = 0; j < 14; j++) s.tri_of[i][j] = -1;
    int nt = 0;
    for (int i = 0; i < 14; i++)
        for (int j = i + 1; j < 14; j++)
            if (dir_adjacent(i, j)) {
                s.tri_i[nt] = (uint8_t)i;
                s.tri_j[nt] = (uint8_t)j;
                s.tri_of[i][j] = (int8_t)nt;
                s.tri_of[j][i] = (int8_t)nt;
                int c = 0;
                for (int k = 0; k < 14; k++)
                    if (k != i && k != j && dir_adjacent(i, k) && dir_adjacent(j, k)) s.tri_k[nt][c++] = (uint8_t)k;
                nt++;
            }
    int nq = 0;
    for (int i = 0; i < 14; i++)
        for (int j = i + 1; j < 14; j++)
            for (int k = j + 1; k < 14; k++)
                if (dir_adjacent(i, j) && dir_adjacent(i, k) && dir_adjacent(j, k)) {
                    s.tet_s[nq][0] = (uint8_t)i;
                    s.tet_s[nq][1] = (uint8_t)j;
                    s.tet_s[nq][2] = (uint8_t)k;
                    s.tet_f[nq][0] = (uint8_t)s.tri_of[i][j];
                    s.tet_f[nq][1] = (uint8_t)s.tri_of[i][k];
                    s.tet_f[nq][2] = (uint8_t)s.tri_of[j][k];
                    nq++;
                }
    for (int t = 0; t < 36; t++)
        for (int c = 0; c < 2; c++) {
            int i = s.tri_i[t], j = s.tri_j[t], k = s.tri_k[t][c];
            int a = i, b = j, d = k;
            // sort ascending
            if (a > b) { int x = a; a = b; b = x; }
            if (b > d) { int x = b; b = d; d = x; }
            if (a > b) { int x = a; a = b; b = x; }
            for (int q = 0; q < 24; q++)
                if (s.tet_s[q][0] == a && s.tet_s[q][1] == b && s.tet_s[q][2] == d) s.tri_q[t][c] = (uint8_t)q;
        }
    for (int i = 0; i < 14; i++)
        for (int c = 0; c < s.adj_n[i]; c++) s.adj_t[i][c] = (uint8_t)s.tri_of[i][s.adj_l[i][c]];
    for (int o = 0; o < 27; o++) s.slot_of[o] = -1;
    for (int k = 0; k < 14; k++) s.slot_of[(DIR[k][0] + 1) * 9 + (DIR[k][1] + 1) * 3 + (DIR[k][2] + 1)] = (int8_t)k;
    return s;
}

constexpr StarTables ST = make_star_tables();

// ---------------------------------------------------------------------------------------------
// Reference simplex IDs
// ---------------------------------------------------------------------------------------------
struct Cell {
    i64 x, y, z;
};

DMTZ_HD i64 vidx(const Dims& d, i64 x, i64 y, i64 z) { return x + y * d.W + z * d.W * d.H; }
DMTZ_HD void vxyz(const Dims& d, i64 v, i64& x, i64& y, i64& z) {
    i64 r;
    divmod(v, d.fW, r, x);
    divmod(r, d.fH, z, y);
}
DMTZ_HD bool inb(const Dims& d, i64 x, i64 y, i64 z) {
    return x >= 0 && x < d.W && y >= 0 && y < d.H && z >= 0 && z < d.D;
}

struct Counts {
    i64 N, E, F, T;
    i64 eoff[8];  // edge block offsets (7 blocks + end)
    i64 foff[5];  // triangle block offsets (4 blocks + end)
};

// Simplex counts main_3d.cu:6401-6415; block layout main_3d.cu:1771-1779 and 439-513.
DMTZ_HD Counts make_counts(const Dims& d) {
    const i64 W = d.W, H = d.H, D = d.D;
    Counts c;
    c.N = W * H * D;
    i64 eb[7] = {(W - 1) * H * D,       W * (H - 1) * D,       W * H * (D - 1),
                 (W - 1) * (H - 1) * D, W * (H - 1) * (D - 1), (W - 1) * H * (D - 1),
                 (W - 1) * (H - 1) * (D - 1)};
    c.eoff[0] = 0;
    for (int k = 0; k < 7; k++) c.eoff[k + 1] = c.eoff[k] + eb[k];
    c.E = c.eoff[7];
    i64 fb[4] = {2 * D * (W - 1) * (H - 1), 2 * W * (H - 1) * (D - 1), 2 * H * (W - 1) * (D - 1),
                 6 * (W - 1) * (H - 1) * (D - 1)};
    c.foff[0] = 0;
    for (int k = 0; k < 4; k++) c.foff[k + 1] = c.foff[k] + fb[k];
    c.F = c.foff[4];
    c.T = 6 * (W - 1) * (H - 1) * (D - 1);
    return c;
}

// Edge block of a direction pair (main_3d.cu:1771-1779): direction k belongs to block k/2.
// Within-block index from the anchor (componentwise minimum of the endpoints).
DMTZ_HD i64 edge_within(const Dims& d, int blk, i64 x, i64 y, i64 z) {
    const i64 W = d.W, H = d.H, D = d.D;
    switch (blk) {
        case 0: return z * (W - 1) * H + y * (W - 1) + x;
        case 1: return z * W * (H - 1) + x * (H - 1) + y;
        case 2: return y * W * (D - 1) + x * (D - 1) + z;
        case 3: return z * (H - 1) * (W - 1) + y * (W - 1) + x;
        case 4: return x * (H - 1) * (D - 1) + y * (D - 1) + z;
        case 5: return y * (W - 1) * (D - 1) + z * (W - 1) + x;
        default: return z * (W - 1) * (H - 1) + y * (W - 1) + x;
    }
}

// Start offsets of edge blocks 1..6 (equal to Counts::eoff[1..6]), computed from the grid sides:
// the values are uniform, so they stay in scalar registers, whereas a runtime index into an
// offset array would force the array into scratch memory.
struct EdgeBlocks {
    i64 s1, s2, s3, s4, s5, s6;
    DMTZ_HD explicit EdgeBlocks(const Dims& d) {
        const i64 W = d.W, H = d.H, D = d.D;
        s1 = (W - 1) * H * D;
        s2 = s1 + W * (H - 1) * D;
        s3 = s2 + W * H * (D - 1);
        s4 = s3 + (W - 1) * (H - 1) * D;
        s5 = s4 + W * (H - 1) * (D - 1);
        s6 = s5 + (W - 1) * H * (D - 1);
    }
    DMTZ_HD i64 start(int blk) const {
        return (blk > 0 ? s1 - 0 : 0) + (blk > 1 ? s2 - s1 : 0) + (blk > 2 ? s3 - s2 : 0) + (blk > 3 ? s4 - s3 : 0) +
               (blk > 4 ? s5 - s4 : 0) + (blk > 5 ? s6 - s5 : 0);
    }
    DMTZ_HD int block_of(i64 e) const {
        return (int)(e >= s1) + (int)(e >= s2) + (int)(e >= s3) + (int)(e >= s4) + (int)(e >= s5) + (int)(e >= s6);
    }
};

// Vertex at cube-local corner k (bit = dx + 2 dy + 4 dz) of the cube anchored at (ax, ay, az).
DMTZ_HD i64 corner_vidx(const Dims& d, i64 ax, i64 ay, i64 az, int k) {
    return vidx(d, ax + (k & 1), ay + ((k >> 1) & 1), az + ((k >> 2) & 1));
}

// Edge between vertex (x,y,z) and its neighbour in direction k.
DMTZ_HD i64 edge_id_dir(const Dims& d, const Counts& c, i64 x, i64 y, i64 z, int k) {
    i64 x1 = x + DIR[k][0], y1 = y + DIR[k][1], z1 = z + DIR[k][2];
    i64 ax = x < x1 ? x : x1, ay = y < y1 ? y : y1, az = z < z1 ? z : z1;
    int blk = k >> 1;
    (void)c;
    return EdgeBlocks(d).start(blk) + edge_within(d, blk, ax, ay, az);
}

// Endpoints of an edge id, larger index first (main_3d.cu:1895-1899). Out-of-range -> -1, -1.
DMTZ_HD void edge_verts(const Dims& d, const Counts& c, i64 e, i64& a, i64& b) {
    if (e < 0 || e >= c.E) {
        a = b = -1;
        return;
    }
    // block = number of block starts at or below e (offsets are non-decreasing)
    const EdgeBlocks eb(d);
    const int blk = eb.block_of(e);
    i64 r = e - eb.start(blk);
    // block layouts as mixed-radix digits (innermost first)
    i64 x = 0, y = 0, z = 0, x1 = 0, y1 = 0, z1 = 0, k = 0;
    switch (blk) {
        case 0: divmod(r, d.fW1, k, x); divmod(k, d.fH, z, y);
                x1 = x + 1; y1 = y; z1 = z; break;
        case 1: divmod(r, d.fH1, k, y); divmod(k, d.fW, z, x);
                x1 = x; y1 = y + 1; z1 = z; break;
        case 2: divmod(r, d.fD1, k, z); divmod(k, d.fW, y, x);
                x1 = x; y1 = y; z1 = z + 1; break;
        case 3: divmod(r, d.fW1, k, x); divmod(k, d.fH1, z, y);
                x1 = x; y1 = y + 1; z1 = z; x = x + 1; break;
        case 4: divmod(r, d.fD1, k, z); divmod(k, d.fH1, x, y);
                x1 = x; y1 = y + 1; z1 = z + 1; break;
        case 5: divmod(r, d.fW1, k, x); divmod(k, d.fD1, y, z);
                x1 = x; y1 = y; z1 = z + 1; x = x + 1; break;
        default: divmod(r, d.fW1, k, x); divmod(k, d.fH1, z, y);
                 x1 = x; y1 = y + 1; z1 = z + 1; x = x + 1; break;
    }
    i64 p = vidx(d, x, y, z), q = vidx(d, x1, y1, z1);
    a = p > q ? p : q;
    b = p > q ? q : p;
}

// Triangle types: (block, cube-local corner mask). Block 0 = xy plane (2 per square), 1 = yz,
// 2 = xz, 3 = cube interior (6 per cube, all through corners 1 and 6).
// Masks from the vertex tables of main_3d.cu:446-454, 472-480, 496-504, 524-556.
constexpr uint8_t TRI_MASK[4][6] = {
    {0x07, 0x0E, 0, 0, 0, 0},                   // xy: {0,1,2}, {1,2,3}
    {0x45, 0x51, 0, 0, 0, 0},                   // yz: {0,2,6}, {0,4,6}
    {0x13, 0x32, 0, 0, 0, 0},                   // xz: {0,1,4}, {1,4,5}
    {0x43, 0x46, 0x4A, 0x52, 0x62, 0xC2}};      // interior: {0,1,6} {1,2,6} {1,3,6} {1,4,6} {1,5,6} {1,6,7}
// Tetrahedron types: cube-local corner masks, main_3d.cu:628-664 / 692-729.
constexpr uint8_t TET_MASK[6] = {0x47, 0x53, 0x4E, 0xCA, 0x72, 0xE2};

// Triangle id from its anchor and corner mask (the triangle is known to be a lattice triangle).
DMTZ_HD i64 tri_id_local(const Dims& d, const Counts& c, i64 ax, i64 ay, i64 az, uint32_t mask) {
    const i64 W = d.W, H = d.H, D = d.D;
    // plane classification by the corner mask
    // xy: all corners with z bit 0 (mask & 0xF0 == 0)
    if ((mask & 0xF0u) == 0) {
        i64 t = (mask == TRI_MASK[0][0]) ? 0 : 1;
        return c.foff[0] + 2 * (az * (W - 1) * (H - 1) + ay * (W - 1) + ax) + t;
    }
    if ((mask & 0xAAu) == 0) {  // all corners with x bit 0: yz plane
        i64 t = (mask == TRI_MASK[1][0]) ? 0 : 1;
        return c.foff[1] + 2 * (ax * (H - 1) * (D - 1) + ay * (D - 1) + az) + t;
    }
    if ((mask & 0xCCu) == 0) {  // all corners with y bit 0: xz plane
        i64 t = (mask == TRI_MASK[2][0]) ? 0 : 1;
        return c.foff[2] + 2 * (ay * (W - 1) * (D - 1) + az * (W - 1) + ax) + t;
    }
    i64 t = 0;
    for (int k = 0; k < 6; k++)
        if (TRI_MASK[3][k] == mask) t = k;
    return c.foff[3] + 6 * (az * (W - 1) * (H - 1) + ay * (W - 1) + ax) + t;
}

DMTZ_HD uint32_t corner_bit(i64 x, i64 y, i64 z, i64 ax, i64 ay, i64 az) {
    return 1u << ((x - ax) + 2 * (y - ay) + 4 * (z - az));
}

DMTZ_HD i64 tri_id3(const Dims& d, const Counts& c, i64 p, i64 q, i64 r) {
    i64 px, py, pz, qx, qy, qz, rx, ry, rz;
    vxyz(d, p, px, py, pz);
    vxyz(d, q, qx, qy, qz);
    vxyz(d, r, rx, ry, rz);
    i64 ax = px < qx ? px : qx; ax = ax < rx ? ax : rx;
    i64 ay = py < qy ? py : qy; ay = ay < ry ? ay : ry;
    i64 az = pz < qz ? pz : qz; az = az < rz ? az : rz;
    uint32_t m = corner_bit(px, py, pz, ax, ay, az) | corner_bit(qx, qy, qz, ax, ay, az) |
                 corner_bit(rx, ry, rz, ax, ay, az);
    return tri_id_local(d, c, ax, ay, az, m);
}

DMTZ_HD i64 tet_id4(const Dims& d, const Counts& c, i64 p, i64 q, i64 r, i64 s) {
    (void)c;
    i64 px, py, pz, qx, qy, qz, rx, ry, rz, sx, sy, sz;
    vxyz(d, p, px, py, pz);
    vxyz(d, q, qx, qy, qz);
    vxyz(d, r, rx, ry, rz);
    vxyz(d, s, sx, sy, sz);
    i64 ax = px < qx ? px : qx; ax = ax < rx ? ax : rx; ax = ax < sx ? ax : sx;
    i64 ay = py < qy ? py : qy; ay = ay < ry ? ay : ry; ay = ay < sy ? ay : sy;
    i64 az = pz < qz ? pz : qz; az = az < rz ? az : rz; az = az < sz ? az : sz;
    uint32_t m = corner_bit(px, py, pz, ax, ay, az) | corner_bit(qx, qy, qz, ax, ay, az) |
                 corner_bit(rx, ry, rz, ax, ay, az) | corner_bit(sx, sy, sz, ax, ay, az);
    i64 t = -1;
    for (int k = 0; k < 6; k++)
        if (TET_MASK[k] == m) t = k;
    if (t < 0) return -1;
    return 6 * (az * (d.W - 1) * (d.H - 1) + ay * (d.W - 1) + ax) + t;
}

DMTZ_HD void sort_desc3(i64& a, i64& b, i64& c) {
    i64 t;
    if (a < b) { t = a; a = b; b = t; }
    if (a < c) { t = a; a = c; c = t; }
    if (b < c) { t = b; b = c; c = t; }
}

// Vertices of a triangle id, descending index order. Decoding follows the block layout of
// main_3d.cu:439-556; an id beyond the last block decodes arithmetically like the reference and
// may give indices >= N (callers drop those).
DMTZ_HD void tri_verts(const Dims& d, const Counts& c, i64 t, i64& v1, i64& v2, i64& v3) {
    i64 ax, ay, az, q;
    uint32_t mask;
    if (t < c.foff[1]) {
        i64 b = t / 2;
        divmod(b, d.fW1, q, ax);
        divmod(q, d.fH1, az, ay);
        mask = TRI_MASK[0][t % 2 != 0 ? 1 : 0];  // a negative odd id takes the odd pattern, like the reference
    } else if (t < c.foff[2]) {
        i64 u = t - c.foff[1], b = u / 2;
        divmod(b, d.fD1, q, az);
        divmod(q, d.fH1, ax, ay);
        mask = TRI_MASK[1][u % 2];
    } else if (t < c.foff[3]) {
        i64 u = t - c.foff[2], b = u / 2;
        divmod(b, d.fW1, q, ax);
        divmod(q, d.fD1, ay, az);
        mask = TRI_MASK[2][u % 2];
    } else {
        i64 u = t - c.foff[3], b = u / 6;
        divmod(b, d.fW1, q, ax);
        divmod(q, d.fH1, az, ay);
        mask = TRI_MASK[3][u % 6];
    }
    // the three set corners of the mask, in increasing corner order
    const int k0 = __builtin_ctz(mask);
    mask &= mask - 1;
    const int k1 = __builtin_ctz(mask);
    mask &= mask - 1;
    const int k2 = __builtin_ctz(mask);
    v1 = corner_vidx(d, ax, ay, az, k0);
    v2 = corner_vidx(d, ax, ay, az, k1);
    v3 = corner_vidx(d, ax, ay, az, k2);
    sort_desc3(v1, v2, v3);
}

DMTZ_HD void sort_desc4(i64& a, i64& b, i64& c, i64& e) {
    i64 t;
    if (a < b) { t = a; a = b; b = t; }
    if (a < c) { t = a; a = c; c = t; }
    if (a < e) { t = a; a = e; e = t; }
    if (b < c) { t = b; b = c; c = t; }
    if (b < e) { t = b; b = e; e = t; }
    if (c < e) { t = c; c = e; e = t; }
}

// Vertices of a tet id, descending index order (main_3d.cu:672-740).
DMTZ_HD void tet_verts(const Dims& d, i64 q, i64& v1, i64& v2, i64& v3, i64& v4) {
    i64 b = q / 6;
    const i64 type = q % 6;
    uint32_t mask = TET_MASK[type < 0 ? 5 : type];  // a negative id takes the last type, like the reference
    i64 az, ay, ax, r;
    divmod(b, d.fW1, r, ax);
    divmod(r, d.fH1, az, ay);
    const int k0 = __builtin_ctz(mask);
    mask &= mask - 1;
    const int k1 = __builtin_ctz(mask);
    mask &= mask - 1;
    const int k2 = __builtin_ctz(mask);
    mask &= mask - 1;
    const int k3 = __builtin_ctz(mask);
    v1 = corner_vidx(d, ax, ay, az, k0);
    v2 = corner_vidx(d, ax, ay, az, k1);
    v3 = corner_vidx(d, ax, ay, az, k2);
    v4 = corner_vidx(d, ax, ay, az, k3);
    sort_desc4(v1, v2, v3, v4);
}

// Slot of vertex u in the star of vertex m (u must be a 14-neighbour of m), else -1.
DMTZ_HD int slot_between(const Dims& d, i64 m, i64 u) {
    i64 mx, my, mz, ux, uy, uz;
    vxyz(d, m, mx, my, mz);
    vxyz(d, u, ux, uy, uz);
    i64 dx = ux - mx, dy = uy - my, dz = uz - mz;
    if (dx < -1 || dx > 1 || dy < -1 || dy > 1 || dz < -1 || dz > 1) return -1;
    return ST.slot_of[(dx + 1) * 9 + (dy + 1) * 3 + (dz + 1)];
}

}  // namespace dmtz
