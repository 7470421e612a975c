// driver.h: device kernels, the context and the host driver of the topology-correction path,
// shared by the library's translation units (dmtz_core.hip: contexts; dmtz_f32.hip and
// dmtz_f64.hip: the entry points of include/dmtz.h per field type). Everything here has internal
// linkage, so each unit compiles the templates it instantiates.
//
// MI355X (gfx950) implementation of the DMTz topology-correction path behind include/dmtz.h.
//
// Device work (one HIP stream per context):
//   k_code        lower star of every vertex -> 32-byte star code (the gradient of
//                 main_3d.cu:1930-2316 restricted to each vertex's lower star)
//   k_critf       f-critical status of the 75 local cells of every vertex (once per run)
//   k_detect      mismatch records from the g star codes and the f-critical masks
//                 (get_false_* main_3d.cu:2492-2550), 4 typed lists
//   k_fix_cp      repair rules over a record list -> edit set (main_3d.cu:2554-4303)
//   k_apply       epsilon-checked edit of the edit set (applyDeltaBuffer1 main_3d.cu:1122-1164)
//   k_dirty, k_code_list
//                 re-evaluation of the g star codes around the vertices whose value changed
//   k_fcrit       f-critical cells (for separatrix checks and filtration)
//   k_ds_geo, k_as_branch + k_as_combine, k_ds_branch + k_ds_combine, k_labels
//                 separatrix checks (main_3d.cu:4437-5062)
//   k_fix_paths   separatrix repairs (main_3d.cu:5135-5295)
//   k_pin         filtration (main_3d.cu:2405-2490)
//   k_zone_cells, k_zone, k_snap_boxes, k_tie_mark, k_tie_raise, k_freeze_all
//                 the optional finisher at stalls (no counterpart in the reference)
// Host: the sweep loop of main_3d.cu:6784-7149; per-sweep counts are read back like the
// reference's cudaMemcpyFromSymbol calls (main_3d.cu:6918-6921, 6989-6992, 7096-7143).
//
// Records: a mismatched (or f-critical) cell is stored as (vertex << 7) | local cell, where the
// vertex is the cell's owner in g (in f for f-critical lists) and the local cell numbers the 75
// cells around a vertex (lower_star.h). Lists are filled with wave-aggregated atomics: the set is
// deterministic, the order is not, and every consumer is order independent.
#pragma once
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <algorithm>
#include <atomic>
#include <thread>
#include <vector>

#include "../../include/dmtz.h"
#include "mismatch.h"
#include "repair.h"

using namespace dmtz;
typedef unsigned long long ull;

namespace {

enum : int {
    K_MIN = 0, K_S1 = 1, K_S2 = 2, K_MAX = 3,  // record lists per cell type
    K_TGT = 4,                                  // edit set size
    K_DS = 5, K_AS = 6,                         // separatrix mismatch lists
    K_MOVED = 7,                                // edits that changed a value or a counter
    K_FC0 = 8,                                  // f-critical lists per type: 8..11
    K_WALK = 12,                                // capped walks
    K_CHG = 13, K_DIRTY = 14,                   // changed vertices, dirty stars
    K_FINISH = 15,                              // finisher tie-repair marks
    K_STARS = 16,                               // star codes re-evaluated from dirty lists
    K_DSW = 17,                                 // 2-saddle wall mismatch list
    K_WOVF = 18,                                // wall walks that outgrew their visited set
    K_ASNEXT = 19,                              // next branch of the work-stealing path walk
    K_DTILES = 20,                              // tiles whose masks the last detection read
    K_NCTR = 21
};

constexpr uint8_t FROZEN = 7;  // counter value of a vertex restored to f by the finisher

constexpr int BLOCK = 256;

// ------------------------------------------------------------------------------------------------
// Wave64 helpers
// ------------------------------------------------------------------------------------------------
__device__ __forceinline__ int wave_excl_scan(int x, int& total) {
    const int lane = __lane_id();
    int inc = x;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        int y = __shfl_up(inc, o, 64);
        if (lane >= o) inc += y;
    }
    total = __shfl(inc, 63, 64);
    return inc - x;
}

// Every lane of the wave must call this. Returns the lane's first slot in the list.
__device__ __forceinline__ ull wave_alloc(ull* ctr, int n) {
    int total;
    int pre = wave_excl_scan(n, total);
    ull base = 0;
    if (__lane_id() == 0 && total > 0) base = atomicAdd(ctr, (ull)total);
    unsigned lo = __shfl((unsigned)(base & 0xffffffffu), 0, 64);
    unsigned hi = __shfl((unsigned)(base >> 32), 0, 64);
    return (((ull)hi << 32) | lo) + (ull)pre;
}

// Every thread of the block must call this (block-uniform control flow). Returns the thread's
// first slot in the list, claimed with one atomic per block: appends of every wave to one
// counter are serialised at the memory side, so a wave-level claim over a list of millions of
// entries costs more than the work that fills it.
__device__ __forceinline__ ull block_alloc(ull* ctr, int n) {
    __shared__ int wsum[BLOCK / 64];
    __shared__ ull bbase;
    int wt;
    const int pre = wave_excl_scan(n, wt);
    const int w = (int)(threadIdx.x >> 6);
    if (__lane_id() == 0) wsum[w] = wt;
    __syncthreads();
    int off = 0, tot = 0;
    for (int i = 0; i < BLOCK / 64; i++) {
        off += i < w ? wsum[i] : 0;
        tot += wsum[i];
    }
    if (threadIdx.x == 0) bbase = tot ? atomicAdd(ctr, (ull)tot) : 0;
    __syncthreads();
    const ull r = bbase + (ull)(off + pre);
    __syncthreads();  // wsum and bbase are reused by the block's next call
    return r;
}

// Every thread of the block must call this. Adds the number of threads with x set to *ctr.
__device__ __forceinline__ void block_add(ull* ctr, bool x) {
    __shared__ int wcnt[BLOCK / 64];
    const ull b = __ballot(x ? 1 : 0);
    if (__lane_id() == 0) wcnt[threadIdx.x >> 6] = __popcll(b);
    __syncthreads();
    if (threadIdx.x == 0) {
        int tot = 0;
        for (int i = 0; i < BLOCK / 64; i++) tot += wcnt[i];
        if (tot) atomicAdd(ctr, (ull)tot);
    }
    __syncthreads();
}

// Every lane of the wave must call this. Adds the number of lanes with x set to *ctr.
__device__ __forceinline__ void wave_add(ull* ctr, bool x) {
    const ull b = __ballot(x ? 1 : 0);
    if (__lane_id() == 0 && b) atomicAdd(ctr, (ull)__popcll(b));
}

struct Lists4 {
    uint64_t* l[4];
    ull cap;
};

// Dense emitters handle EMIT_V vertices per thread: thread t of a block takes vertices
// base + k * BLOCK + t (k < EMIT_V), so each load of a wave is contiguous.
constexpr int EMIT_V = 8;
constexpr i64 EMIT_TILE = (i64)BLOCK * EMIT_V;

// Append the set bits of the masks of a block's EMIT_TILE vertices to the four typed lists. Space
// is claimed with one atomic per type per block after an LDS scan of the per-thread counts;
// blocks with nothing to append return at once. Every thread of the block must call this.
__device__ __forceinline__ void block_emit(i64 base, const uint64_t (&lo)[EMIT_V], const uint32_t (&hi)[EMIT_V],
                                           const Lists4& out, ull* ctr) {
    __shared__ uint32_t wtot[4][BLOCK / 64];
    __shared__ ull wbase[4][BLOCK / 64];
    int cnt[4] = {0, 0, 0, 0};
#pragma unroll
    for (int k = 0; k < EMIT_V; k++) {
        cnt[0] += (int)(lo[k] & 1ull);
        cnt[1] += __popcll((lo[k] >> 1) & 0x3FFFull);
        cnt[2] += __popcll((lo[k] >> 15) & ((1ull << 36) - 1));
        cnt[3] += __popc(hi[k]);
    }
    if (__syncthreads_or(cnt[0] | cnt[1] | cnt[2] | cnt[3]) == 0) return;
    const int wave = threadIdx.x >> 6;
    int pre[4];
#pragma unroll
    for (int t = 0; t < 4; t++) {
        int tot;
        pre[t] = wave_excl_scan(cnt[t], tot);
        if (__lane_id() == 0) wtot[t][wave] = (uint32_t)tot;
    }
    __syncthreads();
    if (threadIdx.x < 4) {
        const int t = threadIdx.x;
        ull sum = 0;
        for (int w = 0; w < BLOCK / 64; w++) sum += wtot[t][w];
        ull b = sum ? atomicAdd(ctr + t, sum) : 0;
        for (int w = 0; w < BLOCK / 64; w++) {
            wbase[t][w] = b;
            b += wtot[t][w];
        }
    }
    __syncthreads();
    ull at[4];
#pragma unroll
    for (int t = 0; t < 4; t++) at[t] = wbase[t][wave] + (ull)pre[t];
#pragma unroll
    for (int k = 0; k < EMIT_V; k++) {
        const uint64_t key = (uint64_t)(base + (i64)k * BLOCK + threadIdx.x) << 7;
        if ((lo[k] & 1ull) && at[0] < out.cap) out.l[0][at[0]] = key | LC_VERT;
        at[0] += lo[k] & 1ull;
        uint64_t m = (lo[k] >> 1) & 0x3FFFull;
        while (m) {
            const int b = __builtin_ctzll(m);
            m &= m - 1;
            if (at[1] < out.cap) out.l[1][at[1]] = key | (uint64_t)(LC_EDGE + b);
            at[1]++;
        }
        m = (lo[k] >> 15) & ((1ull << 36) - 1);
        while (m) {
            const int b = __builtin_ctzll(m);
            m &= m - 1;
            if (at[2] < out.cap) out.l[2][at[2]] = key | (uint64_t)(LC_TRI + b);
            at[2]++;
        }
        uint32_t h = hi[k];
        while (h) {
            const int b = __builtin_ctz(h);
            h &= h - 1;
            if (at[3] < out.cap) out.l[3][at[3]] = key | (uint64_t)(LC_TET + b);
            at[3]++;
        }
    }
}

// ------------------------------------------------------------------------------------------------
// Kernels
// ------------------------------------------------------------------------------------------------
// Star codes of every vertex (dense) or of the listed vertices.
// 75-bit mask over the local cells of a vertex: the f-critical status of every local cell
// (mismatch.h: local_crit), or the mismatched cells a vertex owns in g.
struct __align__(16) CritMask {
    uint64_t lo;
    uint32_t hi, pad;
};

// The lower-star classification is fully unrolled straight-line code of about 5k instructions
// needing about 100 VGPRs (lower_star.h keeps its live state small). CODE_BLOCKS_PER_CU = 4 caps
// the code kernels at 128 VGPRs: four waves per SIMD to hide the loads of the star values.
constexpr int CODE_BLOCKS_PER_CU = 4;

// own (nullable) receives the critical cells each vertex owns, from the lower star itself (the
// first pass of the f-critical masks, without reading the codes back).
template <class T>
__global__ __launch_bounds__(BLOCK, CODE_BLOCKS_PER_CU) void k_code(const T* fld, Dims d, Counts c, StarCode* code,
                                                                    CritMask* own) {
    const i64 v = (i64)blockIdx.x * BLOCK + threadIdx.x;
    if (v >= c.N) return;
    LowerStar ls;
    star_of(fld, d, v, ls);
    encode_star(ls, code[v]);
    if (own) {
        CritMask m;
        crit_mask(ls, m.lo, m.hi);
        m.pad = 0;
        own[v] = m;
    }
}

// g star code of vertex v and its mismatch mask: the cells v owns in g whose critical status in
// g differs from f's (owner_mismatch, main_3d.cu:2492-2550). Returns whether the code changed.
// nz_old / nz_new (nullable) receive whether the mask was nonzero before and after.
template <class T>
__device__ __forceinline__ bool code_g_at(const T* g, Dims d, i64 v, StarCode* code, const CritMask* crit_f,
                                          CritMask* mm, bool* nz_old = nullptr, bool* nz_new = nullptr) {
    LowerStar ls;
    star_of(g, d, v, ls);
    StarCode sc;
    encode_star(ls, sc);
    const StarCode old = code[v];
    code[v] = sc;
    const CritMask cf = crit_f[v];
    uint64_t c_lo, i_lo;  // owner_mismatch of sc, from the lower star's own masks
    uint32_t c_hi, i_hi;
    crit_mask(ls, c_lo, c_hi);
    member_mask(ls, i_lo, i_hi);
    CritMask m;
    m.lo = (c_lo ^ cf.lo) & i_lo;
    m.hi = (c_hi ^ cf.hi) & i_hi;
    m.pad = 0;
    if (nz_old) {
        const CritMask o = mm[v];
        *nz_old = (o.lo | (uint64_t)o.hi) != 0;
    }
    if (nz_new) *nz_new = (m.lo | (uint64_t)m.hi) != 0;
    mm[v] = m;
    return ((old.w[0] ^ sc.w[0]) | (old.w[1] ^ sc.w[1]) | (old.w[2] ^ sc.w[2]) | (old.w[3] ^ sc.w[3])) != 0;
}

// Mark vertex v and the (up to 8) cubes that have it as a corner in the dirty bitmaps of the
// path checks.
__device__ __forceinline__ void mark_vertex(Dims d, uint32_t v, uint32_t* cubes, uint32_t* vbits) {
    atomicOr(vbits + (v >> 5), 1u << (v & 31u));
    i64 x, y, z;
    vxyz(d, (i64)v, x, y, z);
    for (int k = 0; k < 8; k++) {
        const i64 ax = x - (k & 1), ay = y - ((k >> 1) & 1), az = z - ((k >> 2) & 1);
        if (ax < 0 || ay < 0 || az < 0 || ax > d.W - 2 || ay > d.H - 2 || az > d.D - 2) continue;
        const ull cube = (ull)(az * (d.W - 1) * (d.H - 1) + ay * (d.W - 1) + ax);
        atomicOr(cubes + (cube >> 5), 1u << (cube & 31u));
    }
}

// Dense re-evaluation; tile_nz (zeroed before) receives, per EMIT_TILE vertices, the number with
// a nonzero mismatch mask (a wave's 64 vertices lie in one tile).
template <class T>
__global__ __launch_bounds__(BLOCK, CODE_BLOCKS_PER_CU) void k_code_g(const T* g, Dims d, Counts c, StarCode* code,
                                                  const CritMask* crit_f, CritMask* mm, uint32_t* tile_nz) {
    const i64 v = (i64)blockIdx.x * BLOCK + threadIdx.x;
    bool nz = false;
    if (v < c.N) code_g_at(g, d, v, code, crit_f, mm, nullptr, &nz);
    const ull b = __ballot(nz ? 1 : 0);
    if (__lane_id() == 0 && b) atomicAdd(tile_nz + (v / EMIT_TILE), (uint32_t)__popcll(b));
}

// The list length is read on the device (*n_ptr, written by the kernel that filled the list), so
// the host launches it without waiting; the grid strides over the list. Block 0 adds the length
// to *stars, the count of star codes re-evaluated in the run. With cubes != null, every vertex
// whose code changed is marked for the path checks (their results read nothing but star codes).
// tile_nz follows the masks whose nonzero-ness changes.
template <class T>
__global__ __launch_bounds__(BLOCK, CODE_BLOCKS_PER_CU) void k_code_list_g(const T* g, Dims d, StarCode* code,
                                                                         const CritMask* crit_f, CritMask* mm,
                                                                         const uint32_t* list, const ull* n_ptr,
                                                                         ull* stars, uint32_t* cubes,
                                                                         uint32_t* vbits, uint32_t* tile_nz) {
    const ull n = *n_ptr;
    if (blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(stars, n);
    for (ull i = (ull)blockIdx.x * BLOCK + threadIdx.x; i < n; i += (ull)gridDim.x * BLOCK) {
        const uint32_t v = list[i];
        bool nz_old = false, nz_new = false;
        const bool changed = code_g_at(g, d, (i64)v, code, crit_f, mm, &nz_old, &nz_new);
        if (nz_old != nz_new) atomicAdd(tile_nz + (v / EMIT_TILE), nz_new ? 1u : 0xFFFFFFFFu);
        if (changed && cubes) mark_vertex(d, v, cubes, vbits);
    }
}

// f-critical status of the local cells of every vertex (once per run), in two passes: each
// vertex writes the critical cells it owns (k_code's own output, or k_critf_own from the codes),
// then gathers the ones its 14 neighbours own that contain it (k_critf_gather: 15 coalesced
// 16-byte reads and one store per vertex; the reverse, each owner spreading its cells with
// atomics, wrote five times the masks' size to memory and took twice as long).
__global__ __launch_bounds__(BLOCK) void k_critf_own(const StarCode* cf, i64 n, CritMask* out) {
    const i64 v = (i64)blockIdx.x * BLOCK + threadIdx.x;
    if (v >= n) return;
    uint64_t lo, ilo;
    uint32_t hi, ihi;
    code_masks(cf[v], lo, hi, ilo, ihi);
    CritMask m;
    m.lo = lo & ilo;
    m.hi = hi & ihi;
    m.pad = 0;
    out[v] = m;
}

// Critical cells per type (vertex, edge, triangle, tet) over all owners: one atomic per wave and type.
__global__ __launch_bounds__(BLOCK) void k_crit_count(const StarCode* code, i64 n, ull* ctr) {
    const i64 v = (i64)blockIdx.x * BLOCK + threadIdx.x;
    int cnt[4] = {0, 0, 0, 0};
    if (v < n) {
        uint64_t lo, ilo;
        uint32_t hi, ihi;
        code_masks(code[v], lo, hi, ilo, ihi);
        lo &= ilo;
        hi &= ihi;
        cnt[0] = (int)(lo & 1ull);
        cnt[1] = __popcll((lo >> 1) & 0x3FFFull);
        cnt[2] = __popcll(lo >> 15);
        cnt[3] = __popc(hi);
    }
#pragma unroll
    for (int t = 0; t < 4; t++) {
        int tot;
        wave_excl_scan(cnt[t], tot);
        if (__lane_id() == 0 && tot) atomicAdd(ctr + t, (ull)tot);
    }
}

__global__ __launch_bounds__(BLOCK) void k_critf_gather(const CritMask* own, Dims d, i64 n, CritMask* out) {
    const i64 v = (i64)blockIdx.x * BLOCK + threadIdx.x;
    if (v >= n) return;
    uint64_t lo;
    uint32_t hi;
    crit_gather(
        d, v,
        [&](i64 u, uint64_t& a, uint32_t& b) {
            const CritMask m = own[u];
            a = m.lo;
            b = m.hi;
        },
        lo, hi);
    CritMask m;
    m.lo = lo;
    m.hi = hi;
    m.pad = 0;
    out[v] = m;
}

// Dense mismatch detection: every vertex lists the mismatched cells it owns in g, from the
// per-vertex mismatch masks kept with the g star codes (16 bytes per vertex). Waves without a
// mismatch skip the list appends.
// A tile whose count of nonzero masks is zero is skipped without reading its masks; tiles_read
// counts the tiles whose masks were read. A grid of at most DETECT_BLOCKS blocks strides over
// the tiles: block b takes tiles b, b + grid, ..., reads the counts of 64 of them at once and
// visits only the nonzero ones, so a late sweep with a few mismatches does not dispatch a block
// per tile of the grid.
constexpr unsigned DETECT_BLOCKS = 2048;

__global__ __launch_bounds__(BLOCK) void k_detect(const CritMask* mm, i64 n, const uint32_t* tile_nz, Lists4 out,
                                                  ull* ctr, ull* tiles_read) {
    __shared__ uint64_t nz_mask;
    const i64 ntiles = (n + EMIT_TILE - 1) / EMIT_TILE, stride = (i64)gridDim.x;
    for (i64 first = blockIdx.x; first < ntiles; first += 64 * stride) {
        if (threadIdx.x < 64) {
            const i64 tile = first + (i64)threadIdx.x * stride;
            const uint64_t m = __ballot(tile < ntiles && tile_nz[tile] != 0);
            if (threadIdx.x == 0) {
                nz_mask = m;
                if (m) atomicAdd(tiles_read, (ull)__popcll(m));
            }
        }
        __syncthreads();
        uint64_t todo = nz_mask;
        __syncthreads();  // nz_mask is rewritten by the next group
        while (todo) {
            const i64 base = (first + (i64)__builtin_ctzll(todo) * stride) * EMIT_TILE;
            todo &= todo - 1;
            uint64_t lo[EMIT_V];
            uint32_t hi[EMIT_V];
#pragma unroll
            for (int k = 0; k < EMIT_V; k++) {
                const i64 v = base + (i64)k * BLOCK + threadIdx.x;
                lo[k] = 0;
                hi[k] = 0;
                if (v < n) {
                    const CritMask m = mm[v];
                    lo[k] = m.lo;
                    hi[k] = m.hi;
                }
            }
            block_emit(base, lo, hi, out, ctr);
        }
    }
}

// f-critical cells owned by each vertex in f; type_mask selects which types are listed.
__global__ __launch_bounds__(BLOCK) void k_fcrit(const StarCode* cf, Counts c, Lists4 out, ull* ctr, int type_mask) {
    const i64 base = (i64)blockIdx.x * EMIT_TILE;
    uint64_t keep = 0;
    if (type_mask & 1) keep |= 1ull;
    if (type_mask & 2) keep |= 0x3FFFull << 1;
    if (type_mask & 4) keep |= ((1ull << 36) - 1) << 15;
    uint64_t lo[EMIT_V];
    uint32_t hi[EMIT_V];
#pragma unroll
    for (int k = 0; k < EMIT_V; k++) {
        const i64 v = base + (i64)k * BLOCK + threadIdx.x;
        lo[k] = 0;
        hi[k] = 0;
        if (v < c.N) {
            uint64_t ilo;
            uint32_t ihi;
            code_masks(cf[v], lo[k], hi[k], ilo, ihi);
            lo[k] &= keep;
            if (!(type_mask & 8)) hi[k] = 0;
        }
    }
    block_emit(base, lo, hi, out, ctr);
}

struct TargetSink {
    uint32_t* stamp;
    uint32_t epoch;
    uint32_t* list;
    ull* ctr;
    ull cap;
    __device__ void operator()(i64 v) const {
        uint32_t old = atomicExch(stamp + v, epoch);
        if (old != epoch) {
            ull i = atomicAdd(ctr, 1ull);
            if (i < cap) list[i] = (uint32_t)v;
        }
    }
};

// The targets a lane's rule emits, held in registers so that the rule's chain of reads never
// waits on the edit set: the stamp exchanges of all of them are issued together at the end, and
// the wave claims list slots for the new ones with one atomic. No repair rule emits more than
// four; a fifth target would go through the direct sink at once.
struct HoldSink {
    TargetSink direct;
    uint32_t t0, t1, t2, t3;
    int n;
    __device__ __forceinline__ void operator()(i64 v) {
        if (n < 4) {
            t3 = t2;
            t2 = t1;
            t1 = t0;
            t0 = (uint32_t)v;
            n++;
        } else {
            direct(v);
        }
    }
    // Every lane of the wave must call this. Adds the held targets not yet in this epoch's set.
    __device__ __forceinline__ void flush() const {
        const uint32_t e = direct.epoch;
        uint32_t o0 = e, o1 = e, o2 = e, o3 = e;
        if (n > 0) o0 = atomicExch(direct.stamp + t0, e);
        if (n > 1) o1 = atomicExch(direct.stamp + t1, e);
        if (n > 2) o2 = atomicExch(direct.stamp + t2, e);
        if (n > 3) o3 = atomicExch(direct.stamp + t3, e);
        const bool n0 = o0 != e, n1 = o1 != e, n2 = o2 != e, n3 = o3 != e;
        ull at = wave_alloc(direct.ctr, (int)n0 + (int)n1 + (int)n2 + (int)n3);
        const ull cap = direct.cap;
        if (n0) {
            if (at < cap) direct.list[at] = t0;
            at++;
        }
        if (n1) {
            if (at < cap) direct.list[at] = t1;
            at++;
        }
        if (n2) {
            if (at < cap) direct.list[at] = t2;
            at++;
        }
        if (n3 && at < cap) direct.list[at] = t3;
    }
};

// Branch classes of the repair rules: the first split of each rule, i.e. the cell's partner in f
// (an edge with its vertex or a triangle; a triangle with an edge or a tet) or, when f leaves it
// critical, its partner in g, and for the 2-saddle and maximum rules the next split (how the
// partner's remaining vertex lies against the cell in f or g). The class only orders the records,
// and the edit set does not depend on the order; but records of one class take the same
// branches, so a wave of them runs one branch body instead of all of them. It is read from the
// star codes: the record's vertex owns the cell in g, the f-critical masks tell whether f pairs
// it, and the cell's owner in f (its highest vertex in f) holds f's partner.
constexpr int FIX_CLASS_BITS = 4;  // classes 1..12

template <class T, int TYPE>
__global__ __launch_bounds__(BLOCK) void k_fix_class(const uint64_t* recs, ull n, const T* f, const StarCode* cf,
                                                     const T* g, const StarCode* cg, const CritMask* critf, Dims d,
                                                     Counts c, uint16_t* key) {
    const ull i = (ull)blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n) return;
    const uint64_t r = recs[i];
    const i64 v = (i64)(r >> 7);
    const int lc = (int)(r & 127u);
    const CritMask cm = critf[v];
    const bool fcrit = lc < 51 ? ((cm.lo >> lc) & 1ull) != 0 : ((cm.hi >> (lc - 51)) & 1u) != 0;
    const FieldGrad<T> F(f, cf, d, c), G(g, cg, d, c);
    auto nbv = [&](i64 u, int slot) {
        i64 x, y, z;
        vxyz(d, u, x, y, z);
        return vidx(d, x + DIR[slot][0], y + DIR[slot][1], z + DIR[slot][2]);
    };
    // the cell's owner in f and its local index there
    auto f_owner = [&](i64& m, int& lm) {
        m = v;
        lm = lc;
        cell_at_other_vertices(d, v, lc, [&](i64 w, int lw) {
            if (F.above(w, m)) {
                m = w;
                lm = lw;
            }
        });
    };
    int k = 0;
    if (TYPE == 1) {  // fix_saddle main_3d.cu:2588-3037: partner in f, else in g
        const i64 w = nbv(v, lc - LC_EDGE);
        const i64 v1 = v > w ? v : w, v2 = v > w ? w : v;  // endpoints by index
        if (fcrit) {
            const int nib = code_edge(cg[v], lc - LC_EDGE);
            if (nib < 14) {  // triangle with v3: are the endpoints ordered alike against v3 (:2886-2890)
                const i64 v3 = nbv(v, nib);
                k = G.above(v2, v3) == G.above(v1, v3) ? 6 : 7;
            } else {
                k = nib == 14 ? 5 : 8;
            }
        } else {
            i64 m;
            int lm;
            f_owner(m, lm);
            const int nib = code_edge(cf[m], lm - LC_EDGE);
            if (nib < 14) {  // triangle with v3: is v3 below both endpoints in g or in f (:2643-2700)
                const i64 v3 = nbv(m, nib);
                const bool cheap = (G.above(v1, v3) && G.above(v2, v3)) || (F.above(v1, v3) && F.above(v2, v3));
                k = cheap ? 2 : 3;
            } else {
                k = nib == 14 ? 1 : 4;
            }
        }
    } else if (TYPE == 2) {  // fix_2saddle main_3d.cu:3039-3983
        if (fcrit) {  // partner in g, from the g owner v
            const int t = lc - LC_TRI;
            const int cc = code_tri(cg[v], t);
            if (cc < 2) {  // edge (v, n): its partner in f (:3488-3610)
                const int sl = cc == 0 ? ST.tri_i[t] : ST.tri_j[t];
                const i64 w = nbv(v, sl);
                const int nib = F.above(w, v) ? code_edge(cf[w], sl ^ 1) : code_edge(cf[v], sl);
                k = nib < 14 ? 7 : (nib == 14 ? 8 : 9);
            } else if (cc < 4) {  // tet through n4: is n4 below the face in g (:3613-3980)
                const i64 a = nbv(v, ST.tri_i[t]), b = nbv(v, ST.tri_j[t]), v4 = nbv(v, ST.tri_k[t][cc - 2]);
                k = (G.above(a, v4) && G.above(b, v4)) ? 10 : 11;
            } else {
                k = 12;
            }
        } else {  // partner in f, from the f owner m
            i64 m;
            int lm;
            f_owner(m, lm);
            const int t = lm - LC_TRI;
            const int cc = code_tri(cf[m], t);
            const i64 a = nbv(m, ST.tri_i[t]), b = nbv(m, ST.tri_j[t]);
            if (cc < 2) {  // edge (m, e) and the third vertex u (:3057-3223)
                const i64 e = cc == 0 ? a : b, u = cc == 0 ? b : a;
                const i64 v1 = m > e ? m : e, v2 = m > e ? e : m;
                const bool v1e = G.above(v2, u), v2e = G.above(v1, u);
                const bool o1 = F.above(v2, u), o2 = F.above(v1, u);
                k = (v1e && v2e) ? 1 : ((o1 != o2 && v1e != v2e) ? 2 : 3);
            } else if (cc < 4) {  // tet through n4: is n4 below the face in f (:3226-3477)
                const i64 v4 = nbv(m, ST.tri_k[t][cc - 2]);
                k = (F.above(a, v4) && F.above(b, v4)) ? 4 : 5;
            } else {
                k = 6;
            }
        }
    } else if (TYPE == 3) {  // fix_maximum main_3d.cu:3985-4303: is the left-out vertex below the face
        i64 m = v;
        int lm = lc;
        if (!fcrit) f_owner(m, lm);
        const int q = lm - LC_TET;
        const int cc = code_tet(fcrit ? cg[m] : cf[m], q);
        if (cc < 3) {
            const int s0 = ST.tet_s[q][0], s1 = ST.tet_s[q][1], s2 = ST.tet_s[q][2];
            const int out = cc == 0 ? s0 : (cc == 1 ? s1 : s2);
            const i64 w = nbv(m, out);
            const i64 x = nbv(m, cc == 0 ? s1 : s0), y = nbv(m, cc == 2 ? s1 : s2);
            const FieldGrad<T>& X = fcrit ? G : F;
            k = (fcrit ? 3 : 1) + ((X.above(x, w) && X.above(y, w)) ? 0 : 1);
        } else {
            k = fcrit ? 5 : 6;
        }
    }
    key[i] = (uint16_t)k;
}

// Grouping of one kind's records by branch class: a counting sort over the 2^FIX_CLASS_BITS
// classes. Each block of CS_PER_BLOCK records counts its classes (k_class_hist, class-major
// counts, so that their exclusive scan is every (class, block)'s first output slot), then
// scatters its records there (k_class_scatter). Lanes of a wave that share a class claim their
// slots with one LDS atomic. No pass waits on another block's progress, so the sort runs beside
// other kernels without stalling.
constexpr int CS_BINS = 1 << FIX_CLASS_BITS;
constexpr int CS_V = 8;
constexpr ull CS_PER_BLOCK = (ull)BLOCK * CS_V;

// Slot of this lane among the lanes of its wave with the same bin, from a per-bin LDS counter:
// one LDS atomic per distinct bin in the wave. Lanes with bin < 0 take no slot.
__device__ __forceinline__ uint32_t wave_bin_claim(int bin, uint32_t* h) {
    uint32_t slot = 0;
    ull todo = __ballot(bin >= 0);
    while (todo) {
        const int leader = __ffsll((unsigned long long)todo) - 1;
        const int b = __shfl(bin, leader, 64);
        const ull same = __ballot(bin == b);
        uint32_t base = 0;
        if (__lane_id() == leader) base = atomicAdd(h + b, (uint32_t)__popcll(same));
        base = __shfl(base, leader, 64);
        if (bin == b) slot = base + (uint32_t)__popcll(same & ((1ull << __lane_id()) - 1ull));
        todo &= ~same;
    }
    return slot;
}

__global__ __launch_bounds__(BLOCK) void k_class_hist(const uint16_t* key, ull n, ull nblk, uint32_t* hist) {
    __shared__ uint32_t h[CS_BINS];
    if (threadIdx.x < CS_BINS) h[threadIdx.x] = 0;
    __syncthreads();
    const ull base = (ull)blockIdx.x * CS_PER_BLOCK;
    for (int k = 0; k < CS_V; k++) {
        const ull i = base + (ull)k * BLOCK + threadIdx.x;
        (void)wave_bin_claim(i < n ? (int)(key[i] & (CS_BINS - 1)) : -1, h);
    }
    __syncthreads();
    if (threadIdx.x < CS_BINS) hist[(ull)threadIdx.x * nblk + blockIdx.x] = h[threadIdx.x];
}

__global__ __launch_bounds__(BLOCK) void k_class_scatter(const uint16_t* key, const uint64_t* recs, ull n, ull nblk,
                                                         const uint32_t* first, uint64_t* out) {
    __shared__ uint32_t h[CS_BINS];
    if (threadIdx.x < CS_BINS) h[threadIdx.x] = first[(ull)threadIdx.x * nblk + blockIdx.x];
    __syncthreads();
    const ull base = (ull)blockIdx.x * CS_PER_BLOCK;
    for (int k = 0; k < CS_V; k++) {
        const ull i = base + (ull)k * BLOCK + threadIdx.x;
        const int bin = i < n ? (int)(key[i] & (CS_BINS - 1)) : -1;
        const uint32_t slot = wave_bin_claim(bin, h);
        if (bin >= 0) out[slot] = recs[i];
    }
}

// One kernel per cell type, so each holds only its own repair rule (one kernel over all types
// needs twice the registers of the largest rule, and half the waves).
template <class T, int TYPE>
__global__ __launch_bounds__(BLOCK) void k_fix_cp(const uint64_t* recs, ull n, const T* f, const StarCode* cf,
                                                  const T* g, const StarCode* cg, Dims d, Counts c,
                                                  const uint8_t* counter, double eps, double range, int edit_type,
                                                  TargetSink sink) {
    const ull i = (ull)blockIdx.x * BLOCK + threadIdx.x;
    Repair<T, HoldSink> R(f, cf, g, cg, d, c, counter, eps, range, edit_type, HoldSink{sink, 0u, 0u, 0u, 0u, 0});
    if (i < n) {
        const uint64_t r = recs[i];
        const i64 v = (i64)(r >> 7);
        const int lc = (int)(r & 127u);
        const i64 id = local_cell_id(d, c, v, lc);
        if (TYPE == 0) R.fix_minimum(id);
        else if (TYPE == 1) R.fix_saddle(id);
        else if (TYPE == 2) R.fix_2saddle(id);
        else R.fix_maximum(id);
    }
    R.sink.flush();
}

// Zero the counters whose bit is set in mask: one launch where a memset per run of counters
// would take several.
__global__ __launch_bounds__(64) void k_zero_ctrs(ull* ctr, uint32_t mask) {
    const int k = (int)threadIdx.x;
    if (k < K_NCTR && ((mask >> k) & 1u)) ctr[k] = 0;
}
static_assert(K_NCTR <= 32, "k_zero_ctrs takes a 32-bit mask");
constexpr uint32_t ctr_bit(int k) { return 1u << k; }

// Frozen vertices are never edited. moved counts the edits that changed a value or a counter.
// The edit set's size is read on the device (*n_ptr, at most cap); the grid strides over the set
// in block-uniform steps, so every lane takes part in the wave-wide appends.
template <class T>
__global__ __launch_bounds__(BLOCK) void k_apply(const uint32_t* list, const ull* n_ptr, ull cap, const T* f, T* g,
                                                 uint8_t* counter, double eps, double range, int preserve_type,
                                                 int edit_type, uint32_t* changed, ull* chg_ctr, ull* moved) {
    const ull n = *n_ptr < cap ? *n_ptr : cap;
    for (ull base = (ull)blockIdx.x * BLOCK; base < n; base += (ull)gridDim.x * BLOCK) {
        const ull i = base + threadIdx.x;
        bool ch = false, mv = false;
        i64 v = 0;
        if (i < n) {
            v = list[i];
            const int k0 = counter[v];
            if (k0 != FROZEN) {
                const T fv = f[v];
                const T g0 = g[v];
                T gv = g0;
                int k = k0;
                const double dl = edit_type == 1 ? halving_delta(eps, k) : EditArith<T>::delta0(fv, gv, eps);
                if (preserve_type == 0) {
                    if (fabs_d(dl) > 1e-15 * range && k < 5 && EditArith<T>::step(fv, gv, dl, eps)) {
                        k += 1;
                    } else {
                        k = 6;
                        gv = EditArith<T>::snap(fv, eps);
                    }
                } else {
                    k = 6;
                    gv = EditArith<T>::snap(fv, eps);
                }
                g[v] = gv;
                counter[v] = (uint8_t)k;
                ch = !(gv == g0);  // the order (hence every lower star) only sees value comparisons
                mv = ch || k != k0;
            }
        }
        ull at = block_alloc(chg_ctr, ch ? 1 : 0);
        if (ch) changed[at] = (uint32_t)v;
        block_add(moved, mv);
    }
}

// Finisher zone, part 1: the vertices of the first n[k] records of the four mismatch lists, each
// listed once (claimed through its stamp with a fresh epoch).
__global__ __launch_bounds__(BLOCK) void k_zone_cells(Lists4 recs, ull n0, ull n1, ull n2, ull n3, Dims d,
                                                      uint32_t* stamp, uint32_t epoch, uint32_t* list, ull* ctr) {
    const ull i = (ull)blockIdx.x * BLOCK + threadIdx.x;
    i64 vs[4] = {-1, -1, -1, -1};
    if (i < n0 + n1 + n2 + n3) {
        const int k = i < n0 ? 0 : (i < n0 + n1 ? 1 : (i < n0 + n1 + n2 ? 2 : 3));
        const ull j = i - (k > 0 ? n0 : 0) - (k > 1 ? n1 : 0) - (k > 2 ? n2 : 0);
        const uint64_t r = recs.l[k][j];
        const i64 v = (i64)(r >> 7);
        const int lc = (int)(r & 127u);
        i64 x, y, z;
        vxyz(d, v, x, y, z);
        auto nb = [&](int sl) { return vidx(d, x + DIR[sl][0], y + DIR[sl][1], z + DIR[sl][2]); };
        vs[0] = v;
        if (lc >= LC_EDGE && lc < LC_TRI) {
            vs[1] = nb(lc - LC_EDGE);
        } else if (lc >= LC_TRI && lc < LC_TET) {
            vs[1] = nb(ST.tri_i[lc - LC_TRI]);
            vs[2] = nb(ST.tri_j[lc - LC_TRI]);
        } else if (lc >= LC_TET) {
            vs[1] = nb(ST.tet_s[lc - LC_TET][0]);
            vs[2] = nb(ST.tet_s[lc - LC_TET][1]);
            vs[3] = nb(ST.tet_s[lc - LC_TET][2]);
        }
    }
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const bool in = vs[q] >= 0 && atomicExch(stamp + vs[q], epoch) != epoch;
        ull at = wave_alloc(ctr, in ? 1 : 0);
        if (in) list[at] = (uint32_t)vs[q];
    }
}

// Finisher zone, part 2: the vertices targeted by an edit set of epoch in [from, to).
__global__ __launch_bounds__(BLOCK) void k_zone(const uint32_t* stamp, uint32_t from, uint32_t to, i64 n,
                                                uint32_t* list, ull* ctr) {
    const i64 v = (i64)blockIdx.x * BLOCK + threadIdx.x;
    bool in = false;
    if (v < n) {
        const uint32_t s = stamp[v];
        in = s >= from && s < to;
    }
    ull at = wave_alloc(ctr, in ? 1 : 0);
    if (in) list[at] = (uint32_t)v;
}

// The box B: every vertex within Chebyshev distance r of a zone vertex is listed once (claimed
// through its stamp), snapped unless already frozen (g = snap(f): the terminal edit of
// applyDeltaBuffer1, main_3d.cu:1140-1141; the reference keeps a snap of the 14 neighbours
// commented out at main_3d.cu:1144-1151) and frozen. One work item per (zone vertex, offset in
// the (2r+1)^3 box), grid-stride with a wave-uniform trip count.
template <class T>
__global__ __launch_bounds__(BLOCK) void k_snap_boxes(const uint32_t* zone, ull nz, int r, Dims d, const T* f, T* g,
                                                      uint8_t* counter, double eps, uint32_t* stamp, uint32_t epoch,
                                                      uint32_t* box, ull* box_ctr) {
    const ull side = 2 * (ull)r + 1, cube = side * side * side, total = nz * cube;
    for (ull base = (ull)blockIdx.x * BLOCK; base < total; base += (ull)gridDim.x * BLOCK) {
        const ull j = base + threadIdx.x;
        bool in = false;
        i64 u = 0;
        if (j < total) {
            const i64 v = zone[j / cube];
            const ull o = j % cube;
            i64 x, y, z;
            vxyz(d, v, x, y, z);
            const i64 ux = x + (i64)(o % side) - r, uy = y + (i64)((o / side) % side) - r,
                      uz = z + (i64)(o / (side * side)) - r;
            if (inb(d, ux, uy, uz)) {
                u = vidx(d, ux, uy, uz);
                if (atomicExch(stamp + u, epoch) != epoch) {
                    in = true;
                    if (counter[u] != FROZEN) g[u] = EditArith<T>::snap(f[u], eps);
                    counter[u] = FROZEN;
                }
            }
        }
        ull at = wave_alloc(box_ctr, in ? 1 : 0);
        if (in) box[at] = (uint32_t)u;
    }
}

// Tie repair, marking step of one Jacobi round: a box vertex u is marked when a neighbour w has
// g(w) == g(u), f(w) < f(u) and w > u (the index tie-break orders the pair against f).
template <class T>
__global__ __launch_bounds__(BLOCK) void k_tie_mark(const uint32_t* box, ull n, Dims d, const T* f, const T* g,
                                                    uint32_t* stamp, uint32_t epoch, ull* marked) {
    const ull i = (ull)blockIdx.x * BLOCK + threadIdx.x;
    bool mk = false;
    if (i < n) {
        const i64 u = box[i];
        const T fu = f[u], gu = g[u];
        i64 x, y, z;
        vxyz(d, u, x, y, z);
        for (int k = 0; k < 14; k++) {
            const i64 xx = x + DIR[k][0], yy = y + DIR[k][1], zz = z + DIR[k][2];
            if (!inb(d, xx, yy, zz)) continue;
            const i64 w = vidx(d, xx, yy, zz);
            if (g[w] == gu && f[w] < fu && w > u) {
                mk = true;
                break;
            }
        }
        if (mk) stamp[u] = epoch;
    }
    wave_add(marked, mk);
}

// Tie repair, raising step: marked vertices move to the next representable value.
template <class T>
__global__ __launch_bounds__(BLOCK) void k_tie_raise(const uint32_t* box, ull n, T* g, const uint32_t* stamp,
                                                     uint32_t epoch) {
    const ull i = (ull)blockIdx.x * BLOCK + threadIdx.x;
    if (i < n) {
        const i64 u = box[i];
        if (stamp[u] == epoch) g[u] = next_up(g[u]);
    }
}

// Restore g = f and freeze every vertex (counter 7: never edited again).
template <class T>
__global__ __launch_bounds__(BLOCK) void k_freeze_all(const T* f, T* g, uint8_t* counter, i64 n) {
    const i64 v = (i64)blockIdx.x * BLOCK + threadIdx.x;
    if (v < n) {
        g[v] = f[v];
        counter[v] = FROZEN;
    }
}

// Closed stars of the changed vertices: their lower stars must be re-evaluated. One thread per
// (changed vertex, star slot); each vertex enters the dirty list once per epoch.
// The changed count is read on the device (*n_ptr); the grid strides in block-uniform steps.
__global__ __launch_bounds__(BLOCK) void k_dirty(const uint32_t* changed, const ull* n_ptr, Dims d, uint32_t* stamp,
                                                 uint32_t epoch, uint32_t* dirty, ull* dirty_ctr) {
    const ull n = *n_ptr;
    for (ull base = (ull)blockIdx.x * BLOCK; base < 15 * n; base += (ull)gridDim.x * BLOCK) {
        const ull j = base + threadIdx.x;
        bool fresh = false;
        i64 w = 0;
        if (j < 15 * n) {
            const i64 v = changed[j / 15];
            const int slot = (int)(j % 15);
            i64 x, y, z;
            vxyz(d, v, x, y, z);
            bool ok = true;
            w = v;
            if (slot < 14) {
                i64 nx = x + DIR[slot][0], ny = y + DIR[slot][1], nz = z + DIR[slot][2];
                ok = inb(d, nx, ny, nz);
                w = ok ? vidx(d, nx, ny, nz) : 0;
            }
            if (ok) fresh = atomicExch(stamp + w, epoch) != epoch;
        }
        ull at = block_alloc(dirty_ctr, fresh ? 1 : 0);
        if (fresh) dirty[at] = (uint32_t)w;
    }
}

// Dense form of k_dirty for long change lists, where one returning atomic per (changed vertex,
// star slot) costs more than a pass over all stamps: every pair stores the epoch into its star
// vertex's stamp with a plain store (all writers store the same value), then k_dirty_collect
// lists the vertices whose stamp holds the epoch, in vertex order. Same set as k_dirty.
__global__ __launch_bounds__(BLOCK) void k_dirty_mark(const uint32_t* changed, const ull* n_ptr, Dims d,
                                                      uint32_t* stamp, uint32_t epoch) {
    const ull n = *n_ptr;
    for (ull j = (ull)blockIdx.x * BLOCK + threadIdx.x; j < 15 * n; j += (ull)gridDim.x * BLOCK) {
        const i64 v = changed[j / 15];
        const int slot = (int)(j % 15);
        if (slot == 14) {
            stamp[v] = epoch;
            continue;
        }
        i64 x, y, z;
        vxyz(d, v, x, y, z);
        const i64 nx = x + DIR[slot][0], ny = y + DIR[slot][1], nz = z + DIR[slot][2];
        if (inb(d, nx, ny, nz)) stamp[vidx(d, nx, ny, nz)] = epoch;
    }
}

// DIRTY_QUADS groups of four vertices per thread, one 16-byte load of their stamps each; the
// loads of a wave are contiguous. Space in the list is claimed with one atomic per block (a wave
// per block would make the claims on the one counter the cost of the pass).
constexpr int DIRTY_QUADS = 16;
__global__ __launch_bounds__(BLOCK) void k_dirty_collect(const uint32_t* stamp, i64 n, uint32_t epoch,
                                                         uint32_t* dirty, ull* dirty_ctr) {
    __shared__ int wsum[BLOCK / 64];
    __shared__ ull bbase;
    const i64 base = (i64)blockIdx.x * BLOCK * 4 * DIRTY_QUADS;
    uint64_t m = 0;  // bit 4k + j: vertex base + 4 (k BLOCK + threadIdx.x) + j
    for (int k = 0; k < DIRTY_QUADS; k++) {
        const i64 v0 = base + 4 * ((i64)k * BLOCK + threadIdx.x);
        uint64_t q = 0;
        if (v0 + 3 < n) {
            const uint4 u = *reinterpret_cast<const uint4*>(stamp + v0);
            q = (uint64_t)(u.x == epoch) | (uint64_t)(u.y == epoch) << 1 | (uint64_t)(u.z == epoch) << 2 |
                (uint64_t)(u.w == epoch) << 3;
        } else {
            for (int j = 0; j < 4; j++)
                if (v0 + j < n && stamp[v0 + j] == epoch) q |= (uint64_t)1 << j;
        }
        m |= q << (4 * k);
    }
    int wtotal;
    const int pre = wave_excl_scan(__popcll(m), wtotal);
    const int w = (int)(threadIdx.x >> 6);
    if (__lane_id() == 0) wsum[w] = wtotal;
    __syncthreads();
    int off = 0, tot = 0;
    for (int i = 0; i < BLOCK / 64; i++) {
        off += i < w ? wsum[i] : 0;
        tot += wsum[i];
    }
    if (threadIdx.x == 0) bbase = tot ? atomicAdd(dirty_ctr, (ull)tot) : 0;
    __syncthreads();
    ull at = bbase + (ull)(off + pre);
    while (m) {
        const int b = __ffsll((unsigned long long)m) - 1;
        m &= m - 1;
        dirty[at++] = (uint32_t)(base + 4 * ((i64)(b >> 2) * BLOCK + threadIdx.x) + (b & 3));
    }
}

// Filtration main_3d.cu:2405-2490: every vertex of an f-critical cell is pinned to f - eps.
template <class T>
__global__ __launch_bounds__(BLOCK) void k_pin(const uint64_t* recs, ull n, const T* f, T* g, Dims d, Counts c,
                                               double eps) {
    const ull i = (ull)blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n) return;
    const uint64_t r = recs[i];
    const i64 v = (i64)(r >> 7);
    const int lc = (int)(r & 127u);
    i64 x, y, z;
    vxyz(d, v, x, y, z);
    i64 w[4] = {v, -1, -1, -1};
    auto nb = [&](int k) { return vidx(d, x + DIR[k][0], y + DIR[k][1], z + DIR[k][2]); };
    if (lc >= LC_EDGE && lc < LC_TRI) {
        w[1] = nb(lc - LC_EDGE);
    } else if (lc >= LC_TRI && lc < LC_TET) {
        w[1] = nb(ST.tri_i[lc - LC_TRI]);
        w[2] = nb(ST.tri_j[lc - LC_TRI]);
    } else if (lc >= LC_TET) {
        w[1] = nb(ST.tet_s[lc - LC_TET][0]);
        w[2] = nb(ST.tet_s[lc - LC_TET][1]);
        w[3] = nb(ST.tet_s[lc - LC_TET][2]);
    }
    for (int k = 0; k < 4; k++)
        if (w[k] >= 0) g[w[k]] = EditArith<T>::snap(f[w[k]], eps);
}

struct PathList {
    int64_t* id;   // cell id (edge for descending, tet for ascending)
    int64_t* rec;  // recorded vertex / divergent tet
    ull* ctr;
    ull cap;
    // Every thread of the block must call this; threads with has = false add nothing.
    __device__ void push_block(bool has, i64 a, i64 b) const {
        ull at = block_alloc(ctr, has ? 1 : 0);
        if (has && at < cap) {
            id[at] = a;
            rec[at] = b;
        }
    }
};

struct NoSink {
    __device__ void operator()(i64) const {}
};

// The separatrix checks run in two steps: one thread per (start cell, branch) walks and writes
// its result at a fixed position; a combine step per start cell applies the reference's
// "first branch that fails" rule and appends to the list in convergent code.

// get_false_ds_vpath main_3d.cu:4837-4959 over the f-critical edges.
template <class T>
__global__ __launch_bounds__(BLOCK) void k_ds_geo(const uint64_t* fce, ull n, const T* f, const StarCode* cf, const T* g,
                                                  const StarCode* cg, Dims d, Counts c, PathList out) {
    const ull i = (ull)blockIdx.x * BLOCK + threadIdx.x;
    bool has = false;
    i64 e = -1, rec = -1;
    if (i < n) {
        const uint64_t r = fce[i];
        e = local_cell_id(d, c, (i64)(r >> 7), (int)(r & 127u));
        Repair<T, NoSink> R(f, cf, g, cg, d, c, nullptr, 0, 0, 0, NoSink{});
        has = R.ds_geometry(e, rec);
    }
    out.push_block(has, e, rec);
}

// Descending geometry checks over per-edge data computed once per run: endpoints and their f
// pairings (ds_geometry compares only those with g's). pairing -1 is stored as ~0u.
template <class T>
__global__ __launch_bounds__(BLOCK) void k_ds_prepare(const uint64_t* fce, ull n, const T* f, const StarCode* cf,
                                                      Dims d, Counts c, uint4* dsv, uint32_t* dse) {
    const ull i = (ull)blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n) return;
    const uint64_t r = fce[i];
    const i64 e = local_cell_id(d, c, (i64)(r >> 7), (int)(r & 127u));
    FieldGrad<T> F(f, cf, d, c);
    i64 v1, v2;
    edge_verts(d, c, e, v1, v2);
    const i64 p1 = F.pv(v1), p2 = F.pv(v2);
    dsv[i] = make_uint4((uint32_t)v1, (uint32_t)v2, p1 < 0 ? ~0u : (uint32_t)p1, p2 < 0 ? ~0u : (uint32_t)p2);
    dse[i] = (uint32_t)e;
}

// get_false_ds_vpath (main_3d.cu:4837-4959) as ds_geometry does it: an edge is re-evaluated when
// full is set or one of its endpoints had its g lower star re-evaluated since the last check
// (its g pairing depends on nothing else); res keeps each edge's recorded vertex (-1: none).
template <class T>
__global__ __launch_bounds__(BLOCK) void k_ds_check(ull n, const uint4* dsv, const uint32_t* dse, const T* g,
                                                    const StarCode* cg, Dims d, Counts c, const uint32_t* vbits,
                                                    int full, int64_t* res, PathList out) {
    const ull i = (ull)blockIdx.x * BLOCK + threadIdx.x;
    i64 rec = -1, e = -1;
    if (i < n) {
        const uint4 q = dsv[i];
        e = (i64)dse[i];
        const bool redo = full != 0 || ((vbits[q.x >> 5] >> (q.x & 31u)) & 1u) || ((vbits[q.y >> 5] >> (q.y & 31u)) & 1u);
        if (redo) {
            FieldGrad<T> G(g, cg, d, c);
            const i64 p1 = q.z == ~0u ? -1 : (i64)q.z, p2 = q.w == ~0u ? -1 : (i64)q.w;
            const bool b1 = p1 != -1 && p1 != G.pv((i64)q.x);
            const bool b2 = p2 != -1 && p2 != G.pv((i64)q.y);
            rec = b1 ? (i64)q.x : (b2 ? (i64)q.y : -1);
            res[i] = rec;
        } else {
            rec = res[i];
        }
    }
    out.push_block(rec >= 0, e, rec);
}

// Ascending branch walks (is_false main_3d.cu:4437-4497), thread = (f-critical face, cofacet k).
// Connectivity mode (maxlab != null) first compares g's reached maximum with f's label
// (main_3d.cu:4821-4826). res[2*j] = start tet if the branch fails else -1, res[2*j+1] = the
// divergent tet.
// Connectivity mode compares g's maximum with f's label (get_false_as_vpath_connectivity
// main_3d.cu:4761-4835): maxlab holds f's labels per start face; with lab_from_f the start faces
// are g's critical faces and f's label is computed here (the maximum f reaches from the tet when
// f leaves the face critical, -1 otherwise: labels exist only for f's critical faces,
// main_3d.cu:4689-4759).
template <class T>
__global__ __launch_bounds__(BLOCK) void k_as_branch(const uint64_t* fcf, ull n, const int64_t* maxlab, int lab_from_f,
                                                     const T* f, const StarCode* cf, const T* g, const StarCode* cg,
                                                     Dims d, Counts c, int64_t* res, ull* walk_ctr) {
    const ull j = (ull)blockIdx.x * BLOCK + threadIdx.x;
    if (j >= 2 * n) return;
    const ull i = j >> 1;
    const int k = (int)(j & 1);
    const uint64_t r = fcf[i];
    const i64 t = local_cell_id(d, c, (i64)(r >> 7), (int)(r & 127u));
    Repair<T, NoSink> R(f, cf, g, cg, d, c, nullptr, 0, 0, 0, NoSink{});
    i64 q0, q1;
    R.cofacets(t, q0, q1);
    const i64 q = k ? q1 : q0;
    bool ovf = false, found = false;
    i64 dv = -1;
    bool check = q >= 0;
    if (check && maxlab) check = R.paired_max(q, 1, c.N, ovf) != maxlab[2 * i + k];
    if (check && lab_from_f) {
        const i64 lab = R.F.pf(t) == -1 ? R.paired_max(q, 0, c.N, ovf) : -1;
        check = R.paired_max(q, 1, c.N, ovf) != lab;
    }
    if (check) R.as_walk(q, c.N, found, dv, ovf);
    res[2 * j] = found ? q : -1;
    res[2 * j + 1] = dv;
    if (ovf) atomicAdd(walk_ctr, 1ull);
}

// Precomputed ascending paths (geometry mode). A branch j is (f-critical face j / 2, cofacet
// j % 2); its comparison steps (Repair::as_fpath) are fixed for the whole run.
// The first k steps of a branch are kept by the counting walk, so that only longer branches are
// walked a second time to place their steps; k is chosen per run from a memory budget
// (as_keep_steps), between AS_K_MIN and AS_K_MAX.
constexpr int AS_K_MIN = 8, AS_K_MAX = 64;
constexpr unsigned AS_WS_BLOCKS = 2048;  // blocks of the work-stealing walk (eight per CU)
// Beside the repair sweeps (on a second stream) the walk takes half the wave slots: its waves
// run until every branch is walked, so with all slots the sweeps' kernels would wait for it.
constexpr unsigned AS_WS_BLOCKS_SIDE = 1024;
constexpr size_t AS_KEEP_BYTES = (size_t)4 << 30;
struct StoreFirst {
    uint2* p;
    int k;
    DMTZ_HD void operator()(i64 i, i64 cur, i64 t) const {
        if (i < k) p[i] = make_uint2((uint32_t)cur, (uint32_t)t);
    }
};
inline int as_keep_steps(ull nb) {
    const ull k = nb ? (ull)(AS_KEEP_BYTES / sizeof(uint2)) / nb : (ull)AS_K_MAX;
    return (int)(k < (ull)AS_K_MIN ? AS_K_MIN : (k > (ull)AS_K_MAX ? AS_K_MAX : k));
}
struct StoreSteps {
    uint2* p;
    DMTZ_HD void operator()(i64 i, i64 cur, i64 t) const { p[i] = make_uint2((uint32_t)cur, (uint32_t)t); }
};

template <class T>
__device__ i64 branch_tet(const uint64_t* fcf, ull j, const T* f, const StarCode* cf, Dims d, Counts c) {
    const uint64_t r = fcf[j >> 1];
    const i64 t = local_cell_id(d, c, (i64)(r >> 7), (int)(r & 127u));
    Repair<T, NoSink> R(f, cf, f, cf, d, c, nullptr, 0, 0, 0, NoSink{});
    i64 q0, q1;
    R.cofacets(t, q0, q1);
    return (j & 1) ? q1 : q0;
}

template <class T>
__global__ __launch_bounds__(BLOCK) void k_as_count(const uint64_t* fcf, ull nb, const T* f, const StarCode* cf,
                                                    Dims d, Counts c, ull* cnt, uint2* first, int keep, ull* walk_ctr) {
    const ull j = (ull)blockIdx.x * BLOCK + threadIdx.x;
    if (j >= nb) return;
    const i64 q = branch_tet(fcf, j, f, cf, d, c);
    Repair<T, NoSink> R(f, cf, f, cf, d, c, nullptr, 0, 0, 0, NoSink{});
    bool ovf = false;
    cnt[j] = (ull)R.as_fpath(q, c.N, ovf, StoreFirst{first + j * (ull)keep, keep});
    if (ovf) atomicAdd(walk_ctr, 1ull);
}

// k_as_count with work stealing: path lengths are heavy-tailed, so instead of one branch per
// thread, every lane takes a new branch whenever its walk ends, and the waves run until the
// branches are exhausted. Each loop round every busy lane advances its walk by one step, so a
// wave's lanes stay busy until the last branches. A wave claims AS_WS_CHUNK branches at a time
// from the shared counter (*next) and hands them to its idle lanes, so the one counter sees an
// atomic per chunk rather than one per round. Same per-branch results as k_as_count.
constexpr ull AS_WS_CHUNK = 1024;
template <class T>
__global__ __launch_bounds__(BLOCK) void k_as_count_ws(const uint64_t* fcf, ull nb, const T* f, const StarCode* cf,
                                                       Dims d, Counts c, ull* cnt, uint2* first, int keep, ull* next,
                                                       ull* walk_ctr) {
    Repair<T, NoSink> R(f, cf, f, cf, d, c, nullptr, 0, 0, 0, NoSink{});
    const int lane = __lane_id();
    bool busy = false;
    ull j = 0;
    i64 cur = -1, n = 0;
    ull cb = 0, ce = 0;  // the wave's unclaimed branches [cb, ce) (wave-uniform)
    bool dry = false;    // the shared counter has passed nb (wave-uniform)
    for (;;) {
        const ull idle = __ballot(!busy);
        if (idle && !(dry && cb >= ce)) {
            const ull nid = (ull)__popcll(idle);
            const ull rank = (ull)__popcll(idle & ((1ull << lane) - 1ull));
            const ull avail = ce - cb;
            ull jj = nb;  // this lane's new branch, if idle (nb: none)
            if (nid <= avail) {
                jj = cb + rank;
                cb += nid;
            } else {
                ull nc = 0;
                if (!dry) {
                    const int leader = __ffsll((unsigned long long)idle) - 1;
                    if (lane == leader) nc = atomicAdd(next, AS_WS_CHUNK);
                    const unsigned lo = __shfl((unsigned)(nc & 0xffffffffu), leader, 64);
                    const unsigned hi = __shfl((unsigned)(nc >> 32), leader, 64);
                    nc = ((ull)hi << 32) | lo;
                    dry = nc >= nb;
                }
                if (rank < avail) jj = cb + rank;
                else if (!dry) jj = nc + (rank - avail);
                if (!dry) {
                    ce = nc + AS_WS_CHUNK < nb ? nc + AS_WS_CHUNK : nb;
                    cb = nc + (nid - avail) < ce ? nc + (nid - avail) : ce;
                } else {
                    cb = ce;
                }
            }
            if (!busy && jj < nb) {
                j = jj;
                cur = branch_tet(fcf, j, f, cf, d, c);
                n = 0;
                if (cur < 0) cnt[j] = 0;  // no cofacet on this side: an empty path
                else busy = true;
            }
        }
        if (__ballot(busy) == 0 && dry && cb >= ce) break;
        if (busy) {
            bool fin = R.as_fpath_step(cur, n, StoreFirst{first + j * (ull)keep, keep});
            if (!fin && n > c.N) {  // as_fpath's cap
                atomicAdd(walk_ctr, 1ull);
                fin = true;
            }
            if (fin) {
                cnt[j] = (ull)n;
                busy = false;
            }
        }
    }
}

// Steps of thread t's branch at off[t]: copied from the counting walk's first keep steps, or
// walked again. Thread t takes branch order[t] (the check order of k_len_scatter: a wave's walks
// are alike; off then holds the offsets in that order) or branch t when order is null.
template <class T>
__global__ __launch_bounds__(BLOCK) void k_as_fill(const uint64_t* fcf, ull nb, const uint32_t* order, const T* f,
                                                   const StarCode* cf, Dims d, Counts c, const ull* off,
                                                   const uint2* first, int keep, uint2* steps) {
    const ull t = (ull)blockIdx.x * BLOCK + threadIdx.x;
    if (t >= nb) return;
    const ull j = order ? (ull)order[t] : t;
    const ull a = off[t], n = off[t + 1] - a;
    if (n <= (ull)keep) {
        for (ull k = 0; k < n; k++) steps[a + k] = first[j * (ull)keep + k];
        return;
    }
    const i64 q = branch_tet(fcf, j, f, cf, d, c);
    Repair<T, NoSink> R(f, cf, f, cf, d, c, nullptr, 0, 0, 0, NoSink{});
    bool ovf = false;
    R.as_fpath(q, c.N, ovf, StoreSteps{steps + a});
}

// is_false (main_3d.cu:4437-4497) over the precomputed steps: the first step whose g partner
// face differs from f's. A branch keeps its previous result unless full is set or one of its
// tets lies in a cube marked dirty since the last check (a tet's g pairing can only change when
// one of its vertices had its lower star re-evaluated).
template <class T>
__global__ __launch_bounds__(BLOCK) void k_as_check(ull nb, const ull* off, const uint2* steps, const T* g,
                                                    const StarCode* cg, Dims d, Counts c, const uint32_t* cubes,
                                                    int full, int64_t* res) {
    const ull j = (ull)blockIdx.x * BLOCK + threadIdx.x;
    if (j >= nb) return;
    const ull a = off[j], b = off[j + 1];
    bool redo = full != 0;
    for (ull s = a; s < b && !redo; s++) {
        const uint32_t cube = steps[s].x / 6u;
        redo = ((cubes[cube >> 5] >> (cube & 31u)) & 1u) != 0;
    }
    if (!redo) return;
    FieldGrad<T> G(g, cg, d, c);
    i64 found = -1, dv = -1;
    bool done = a >= b;
    ull s = a;
    while (!done) {
        const uint2 st = steps[s];
        if (G.pt((i64)st.x) != (i64)st.y) {
            found = (i64)steps[a].x;
            dv = (i64)st.x;
        }
        s++;
        done = found >= 0 || s >= b;
    }
    res[2 * j] = found;
    res[2 * j + 1] = dv;
}

// Compact copies for the dirty test of the checks below: 32-bit branch offsets and the cube of
// every step (needs fewer than 2^32 steps). Thread per branch.
__global__ __launch_bounds__(BLOCK) void k_as_compact(ull nb, const ull* off, const uint2* steps, uint32_t* off32,
                                                      uint32_t* cube) {
    const ull j = (ull)blockIdx.x * BLOCK + threadIdx.x;
    if (j >= nb) return;
    off32[j] = (uint32_t)off[j];
    if (j == nb - 1) off32[nb] = (uint32_t)off[nb];
    for (ull s = off[j]; s < off[j + 1]; s++) cube[s] = steps[s].x / 6u;
}

// Step counts in check order (cnt of branch order[t] at t), and a zero at nb for the scan.
__global__ __launch_bounds__(BLOCK) void k_as_gather_cnt(ull nb, const uint32_t* order, const ull* cnt, ull* out) {
    const ull t = (ull)blockIdx.x * BLOCK + threadIdx.x;
    if (t < nb) out[t] = cnt[order[t]];
    else if (t == nb) out[nb] = 0;
}

// Branch order of the checks: the branches without a step first (the checks skip them), then
// the others in chunks of LEN_CHUNK_BLOCKS * CS_PER_BLOCK consecutive branches, within a chunk
// grouped by step count (capped at 255) so that a wave's walks have similar lengths. The
// f-critical faces are listed in about vertex order, so a chunk covers a few planes of the grid
// and the checks read one region at a time: neighbouring branches share the values and star
// codes they read while these are cached. (Grouped by length over the whole grid, most branches
// being a few steps long, every region was read again for each length.)
// A counting sort like k_class_hist / k_class_scatter, with the key read from the counts.
constexpr int LEN_BINS = 256;
constexpr ull LEN_CHUNK_BLOCKS = 64;
__device__ __forceinline__ int len_bin(const ull* cnt, ull j) { return (int)(cnt[j] < 255 ? cnt[j] : 255); }

// Position of (block blk, length bin) in the count table of nblk blocks: bin 0 of every block,
// then the chunks in order, each bin-major (the last chunk may be shorter). Bin 1 of block 0
// is at nblk, so the scan there is the number of branches without a step.
__device__ __forceinline__ ull len_slot(ull blk, int bin, ull nblk) {
    if (bin == 0) return blk;
    const ull b0 = blk / LEN_CHUNK_BLOCKS * LEN_CHUNK_BLOCKS;
    const ull w = nblk - b0 < LEN_CHUNK_BLOCKS ? nblk - b0 : LEN_CHUNK_BLOCKS;
    return nblk + b0 * (LEN_BINS - 1) + (ull)(bin - 1) * w + (blk - b0);
}

__global__ __launch_bounds__(BLOCK) void k_len_hist(ull nb, const ull* cnt, ull nblk, uint32_t* hist) {
    __shared__ uint32_t h[LEN_BINS];
    h[threadIdx.x] = 0;  // BLOCK == LEN_BINS
    __syncthreads();
    const ull base = (ull)blockIdx.x * CS_PER_BLOCK;
    for (int k = 0; k < CS_V; k++) {
        const ull j = base + (ull)k * BLOCK + threadIdx.x;
        (void)wave_bin_claim(j < nb ? len_bin(cnt, j) : -1, h);
    }
    __syncthreads();
    hist[len_slot(blockIdx.x, threadIdx.x, nblk)] = h[threadIdx.x];
}

__global__ __launch_bounds__(BLOCK) void k_len_scatter(ull nb, const ull* cnt, ull nblk, const uint32_t* first,
                                                       uint32_t* order) {
    __shared__ uint32_t h[LEN_BINS];
    h[threadIdx.x] = first[len_slot(blockIdx.x, threadIdx.x, nblk)];
    __syncthreads();
    const ull base = (ull)blockIdx.x * CS_PER_BLOCK;
    for (int k = 0; k < CS_V; k++) {
        const ull j = base + (ull)k * BLOCK + threadIdx.x;
        const int bin = j < nb ? len_bin(cnt, j) : -1;
        const uint32_t slot = wave_bin_claim(bin, h);
        if (bin >= 0) order[slot] = (uint32_t)j;
    }
}
static_assert(BLOCK == LEN_BINS, "k_len_hist keeps one bin per thread");

// Per-step dirty bits of the compact path copies: bit s of sd is set when the tet of step s
// lies in a marked cube. Each wave covers 4 x 64 consecutive steps, lane l the steps
// 64k + l (k = 0..3): the cube ids are read coalesced, 4 bytes per step, with four independent
// loads in flight per lane, and each 64-step word is one ballot.
constexpr int SD_WORDS_PER_WAVE = 4;

__global__ __launch_bounds__(BLOCK) void k_as_step_dirty(ull total, const uint32_t* cube, const uint32_t* cubes,
                                                         uint64_t* sd) {
    const ull w = (ull)blockIdx.x * (BLOCK / 64) + (threadIdx.x >> 6);
    const int lane = (int)(threadIdx.x & 63u);
    const ull base = w * (64 * SD_WORDS_PER_WAVE);
    uint32_t q[SD_WORDS_PER_WAVE];
#pragma unroll
    for (int k = 0; k < SD_WORDS_PER_WAVE; k++) {
        const ull s = base + 64 * k + lane;
        q[k] = s < total ? cube[s] : 0u;
    }
#pragma unroll
    for (int k = 0; k < SD_WORDS_PER_WAVE; k++) {
        const ull s = base + 64 * k + lane;
        const bool dq = s < total && ((cubes[q[k] >> 5] >> (q[k] & 31u)) & 1u) != 0;
        const uint64_t m = __ballot(dq);
        if (lane == 0 && base + 64 * k < total) sd[(base >> 6) + k] = m;
    }
}

// First step in [s, b) whose dirty bit is set, or b.
__device__ __forceinline__ uint32_t next_dirty(const uint64_t* sd, uint32_t s, uint32_t b) {
    uint32_t r = b;
    for (uint32_t w = s >> 6; s < b && w <= (b - 1) >> 6 && r == b; w++) {
        const uint32_t lo = w << 6;
        uint64_t m = sd[w];
        if (lo < s) m &= ~0ull << (s - lo);
        if (b - lo < 64) m &= (1ull << (b - lo)) - 1ull;
        if (m) r = lo + (uint32_t)__builtin_ctzll(m);
    }
    return r;
}

// k_as_check over the compact copies. Thread t checks branch order[t], whose steps are at
// [off32[t], off32[t + 1]) (the copies are laid out in check order, so a wave's offsets and dirty
// words are adjacent); results stay indexed by branch. Threads start at check position t0: the
// positions before it hold the branches without a step, whose result stays empty. Unless full, a
// branch with no dirty step keeps its last result, and a branch whose
// last result was "no mismatch" compares only its dirty steps, in order: the others matched then
// and read unchanged star codes now. A branch that had a mismatch and has a dirty step is walked
// again from its start.
template <class T>
__global__ __launch_bounds__(BLOCK) void k_as_check32(ull t0, ull nb, const uint32_t* order, const uint32_t* off32,
                                                      const uint64_t* sd, const uint2* steps, const T* g,
                                                      const StarCode* cg, Dims d, Counts c, int full, int64_t* res,
                                                      uint32_t* fails) {
    const ull t = t0 + (ull)blockIdx.x * BLOCK + threadIdx.x;
    if (t >= nb) return;
    const uint32_t a = off32[t], b = off32[t + 1];
    uint32_t s = a;
    bool sparse = false;  // visit dirty steps only
    if (!full) {
        s = next_dirty(sd, a, b);
        if (s >= b) return;
    }
    const ull j = order[t];
    if (!full) {
        sparse = res[2 * j] < 0;
        if (!sparse) s = a;
    }
    FieldGrad<T> G(g, cg, d, c);
    i64 found = -1, dv = -1;
    while (s < b && found < 0) {
        const uint2 st = steps[s];
        if (G.pt((i64)st.x) != (i64)st.y) {
            found = (i64)steps[a].x;
            dv = (i64)st.x;
        }
        s = sparse ? next_dirty(sd, s + 1, b) : s + 1;
    }
    // fails bit j mirrors res[2 * j] >= 0 (zeroed before a full check)
    const bool was = !full && !sparse, now = found >= 0;
    if (was != now) atomicXor(fails + (j >> 5), 1u << (j & 31u));
    res[2 * j] = found;
    res[2 * j + 1] = dv;
}

// With fails != null (bit j set when branch j failed), faces with neither branch failing skip the
// result reads.
__global__ __launch_bounds__(BLOCK) void k_as_combine(const int64_t* res, ull n, PathList out, const uint32_t* fails) {
    const ull i = (ull)blockIdx.x * BLOCK + threadIdx.x;
    bool has = false;
    i64 a = -1, b = -1;
    if (i < n && (!fails || ((fails[(2 * i) >> 5] >> ((2 * i) & 31u)) & 3u) != 0)) {
        const int64_t* p = res + 4 * i;
        has = p[0] >= 0 || p[2] >= 0;
        a = p[0] >= 0 ? p[0] : p[2];
        b = p[0] >= 0 ? p[1] : p[3];
    }
    out.push_block(has, a, b);
}

// k_as_combine over the failure bits (bit j set when branch j failed; the bits mirror
// res[2 * j] >= 0): one thread per 32 faces, the 64 branch bits of one word pair, so that a face
// without a failing branch costs one bit instead of a thread. Faces whose bits are set are read
// twice: once to count what the thread appends, once to append it.
__global__ __launch_bounds__(BLOCK) void k_as_combine_bits(const int64_t* res, ull n, PathList out,
                                                           const uint32_t* fails) {
    const ull t = (ull)blockIdx.x * BLOCK + threadIdx.x;
    const ull f0 = t * 32;
    uint64_t fm = 0;  // bit 2i: face f0 + i has a failing branch
    if (f0 < n) {
        uint64_t bits = (uint64_t)fails[2 * t] | ((uint64_t)fails[2 * t + 1] << 32);
        const ull nf = n - f0 < 32 ? n - f0 : 32;
        if (nf < 32) bits &= (1ull << (2 * nf)) - 1ull;
        fm = (bits | (bits >> 1)) & 0x5555555555555555ull;
    }
    uint64_t hm = 0;  // faces that append
    for (uint64_t m = fm; m; m &= m - 1) {
        const int b = __builtin_ctzll(m);
        const int64_t* p = res + 4 * (f0 + (ull)(b >> 1));
        if (p[0] >= 0 || p[2] >= 0) hm |= 1ull << b;
    }
    ull at = block_alloc(out.ctr, __popcll(hm));
    for (uint64_t m = hm; m; m &= m - 1) {
        const int b = __builtin_ctzll(m);
        const int64_t* p = res + 4 * (f0 + (ull)(b >> 1));
        if (at < out.cap) {
            out.id[at] = p[0] >= 0 ? p[0] : p[2];
            out.rec[at] = p[0] >= 0 ? p[1] : p[3];
        }
        at++;
    }
}

// Endpoint labels of f (get_orginal_saddle_min_pairs / _max_pairs, main_3d.cu:4613-4759), one
// thread per (f-critical edge, endpoint) and per (f-critical face, cofacet); ascending labels
// are indexed by face (resolution of the shadowed index at main_3d.cu:4750).
template <class T>
__global__ __launch_bounds__(BLOCK) void k_labels(const uint64_t* fce, ull ne, const uint64_t* fcf, ull nf, const T* f,
                                                  const StarCode* cf, Dims d, Counts c, int64_t* minlab,
                                                  int64_t* maxlab, ull* walk_ctr) {
    const ull j = (ull)blockIdx.x * BLOCK + threadIdx.x;
    Repair<T, NoSink> R(f, cf, f, cf, d, c, nullptr, 0, 0, 0, NoSink{});
    bool ovf = false;
    if (j < 2 * ne) {
        const uint64_t r = fce[j >> 1];
        const i64 e = local_cell_id(d, c, (i64)(r >> 7), (int)(r & 127u));
        i64 a, b;
        edge_verts(d, c, e, a, b);
        minlab[j] = R.f_min_label((j & 1) ? b : a, c.N, ovf);
    } else if (j < 2 * ne + 2 * nf) {
        const ull jj = j - 2 * ne;
        const uint64_t r = fcf[jj >> 1];
        const i64 t = local_cell_id(d, c, (i64)(r >> 7), (int)(r & 127u));
        i64 q0, q1;
        R.cofacets(t, q0, q1);
        const i64 q = (jj & 1) ? q1 : q0;
        maxlab[jj] = q >= 0 ? R.paired_max(q, 0, c.N, ovf) : -1;
    }
    if (ovf) atomicAdd(walk_ctr, 1ull);
}

// get_false_ds_vpath_connectivity main_3d.cu:4961-5062: thread = (f-critical edge, endpoint)
// walks g's descending path; res[4*i + 2*k] = reached g minimum, res[4*i + 2*k + 1] = first
// vertex on the walk whose f and g pairings differ.
template <class T>
__global__ __launch_bounds__(BLOCK) void k_ds_branch(const uint64_t* fce, ull n, const T* f, const StarCode* cf,
                                                     const T* g, const StarCode* cg, Dims d, Counts c, int64_t* res,
                                                     ull* walk_ctr) {
    const ull j = (ull)blockIdx.x * BLOCK + threadIdx.x;
    if (j >= 2 * n) return;
    const uint64_t r = fce[j >> 1];
    const i64 e = local_cell_id(d, c, (i64)(r >> 7), (int)(r & 127u));
    Repair<T, NoSink> R(f, cf, g, cg, d, c, nullptr, 0, 0, 0, NoSink{});
    i64 a, b;
    edge_verts(d, c, e, a, b);
    i64 trouble = -1;
    bool ovf = false;
    res[2 * j] = R.g_min_walk((j & 1) ? b : a, trouble, c.N, ovf);
    res[2 * j + 1] = trouble;
    if (ovf) atomicAdd(walk_ctr, 1ull);
}

// The reference carries trouble_maker from the first endpoint's walk into the second.
__global__ __launch_bounds__(BLOCK) void k_ds_combine(const uint64_t* fce, ull n, const int64_t* res,
                                                      const int64_t* minlab, Dims d, Counts c, PathList out) {
    const ull i = (ull)blockIdx.x * BLOCK + threadIdx.x;
    bool has = false;
    i64 e = -1, rec = -1;
    if (i < n) {
        const uint64_t r = fce[i];
        e = local_cell_id(d, c, (i64)(r >> 7), (int)(r & 127u));
        const int64_t* p = res + 4 * i;
        bool bad1 = p[0] != minlab[2 * i];
        bool bad2 = p[2] != minlab[2 * i + 1];
        has = bad1 || bad2;
        rec = bad1 ? p[1] : (p[1] != -1 ? p[1] : p[3]);
    }
    out.push_block(has, e, rec);
}

// 2-saddle walls, get_false_ds_wall main_3d.cu:5064-5133, made deterministic: from each
// f-critical face, a depth-first walk over f's face -> edge -> paired-face steps with a stack and
// a visited set of its own (the reference shares one stack between threads, defects D2-D5).
// A face is marked when popped; its edges are taken in the order (v1,v2), (v1,v3), (v2,v3) and
// pushed in that order, so the last is popped first. The first edge met whose f partner exists
// and differs from its g partner is recorded as (face, edge), and the walk stops.
//
// Thread t of the launch walks start first + t (or sel[first + t]) with its own part of scratch
// (2 * slots + 2 entries: the visited set and the stack of Repair::wall_walk). A walk that
// outgrows its set is listed in ovf, to be walked again with more slots.
template <class T>
__global__ __launch_bounds__(BLOCK) void k_ds_wall(const uint64_t* starts, const uint32_t* sel, ull first, ull n,
                                                   const T* f, const StarCode* cf, const T* g, const StarCode* cg,
                                                   Dims d, Counts c, uint32_t* scratch, uint32_t slots, PathList out,
                                                   uint32_t* ovf, ull* ovf_ctr) {
    const ull t = (ull)blockIdx.x * BLOCK + threadIdx.x;
    if (first + t >= n) return;
    const ull s = sel ? (ull)sel[first + t] : first + t;
    const uint64_t r = starts[s];
    const i64 face = local_cell_id(d, c, (i64)(r >> 7), (int)(r & 127u));
    Repair<T, NoSink> R(f, cf, g, cg, d, c, nullptr, 0, 0, 0, NoSink{});
    uint32_t* set = scratch + (size_t)t * (2 * (size_t)slots + 2);
    i64 found = -1;
    const bool overflow = R.wall_walk(face, set, set + slots, slots, found);
    if (overflow) {
        ovf[atomicAdd(ovf_ctr, 1ull)] = (uint32_t)s;
    } else if (found >= 0) {
        const ull at = atomicAdd(out.ctr, 1ull);
        if (at < out.cap) {
            out.id[at] = face;
            out.rec[at] = found;
        }
    }
}

// fix_as_vpath / fix_ds_vpath / fix_ds_wall main_3d.cu:5135-5413
template <class T>
__global__ __launch_bounds__(BLOCK) void k_fix_paths(const int64_t* as_rec, ull n_as, const int64_t* ds_rec, ull n_ds,
                                                     const int64_t* wall_rec, ull n_wall, const T* f,
                                                     const StarCode* cf, const T* g, const StarCode* cg, Dims d,
                                                     Counts c, const uint8_t* counter, double eps, double range,
                                                     int edit_type, TargetSink sink) {
    const ull i = (ull)blockIdx.x * BLOCK + threadIdx.x;
    Repair<T, HoldSink> R(f, cf, g, cg, d, c, counter, eps, range, edit_type, HoldSink{sink, 0u, 0u, 0u, 0u, 0});
    if (i < n_as) R.fix_as_vpath(as_rec[i]);
    else if (i < n_as + n_ds) R.fix_ds_vpath(ds_rec[i - n_as]);
    else if (i < n_as + n_ds + n_wall) R.fix_ds_wall(wall_rec[i - n_as - n_ds]);
    R.sink.flush();
}

// Dense pairing arrays for tests (dmtz_gradient_*): one thread per cell of each dimension.
template <class T>
__global__ __launch_bounds__(BLOCK) void k_pairing(const T* fld, const StarCode* code, Dims d, Counts c, int dim, ull n,
                                                   int64_t* out) {
    const ull i = (ull)blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n) return;
    FieldGrad<T> X(fld, code, d, c);
    switch (dim) {
        case 0: out[i] = X.pv((i64)i); break;
        case 1: out[i] = X.pe((i64)i); break;
        case 2: out[i] = X.pf((i64)i); break;
        default: out[i] = X.pt((i64)i); break;
    }
}

inline unsigned nblocks(ull n) { return (unsigned)((n + BLOCK - 1) / BLOCK); }

// Grid of a kernel that reads its element count on the device and strides over it: sized for
// the host's estimate of the count, at least one block and at most STRIDE_BLOCKS (eight blocks
// per CU of a 256-CU device).
constexpr ull STRIDE_BLOCKS = 2048;
inline unsigned stride_blocks(ull estimate) {
    const ull b = nblocks(estimate);
    return (unsigned)(b < 1 ? 1 : (b > STRIDE_BLOCKS ? STRIDE_BLOCKS : b));
}

// The dirty stars are collected by one pass over all N stamps (k_dirty_mark + k_dirty_collect)
// once the (changed vertex, star slot) pairs reach N / DIRTY_DENSE, and by one returning atomic
// per pair (k_dirty) below that.
constexpr ull DIRTY_DENSE = 256;

}  // namespace

// ================================================================================================
// Context
// ================================================================================================
struct dmtz_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    // device buffers
    void* d_f = nullptr;
    void* d_g = nullptr;
    size_t field_bytes = 0;
    uint8_t* d_counter = nullptr;
    uint32_t* d_stamp = nullptr;    // edit-set dedup stamps
    uint32_t* d_dstamp = nullptr;   // dirty-star dedup stamps
    uint32_t* d_targets = nullptr;  // edit set
    uint32_t* d_changed = nullptr;  // edited vertices whose value changed
    uint32_t* d_dirty = nullptr;    // vertices whose g star code must be re-evaluated
    StarCode* d_code_f = nullptr;   // star codes of f and g
    StarCode* d_code_g = nullptr;
    CritMask* d_critf = nullptr;    // f-critical status of every vertex's local cells
    CritMask* d_mm = nullptr;       // mismatched cells each vertex owns in g
    uint32_t* d_tile_nz = nullptr;  // per EMIT_TILE vertices: how many have a nonzero d_mm
    uint64_t* d_rec = nullptr;    // 4 mismatch record lists of rec_cap entries
    uint64_t* d_fc = nullptr;     // 4 f-critical lists of fc_cap entries
    int64_t* d_path = nullptr;    // ds id/rec, as id/rec, wall id/rec (fc_cap each)
    // Repairs of record types 1..3: class keys, grouped records and counting-sort scratch per type,
    // and the second stream of the repair phase.
    bool fix_concurrent = true;  // off in batch lanes: no second stream (repairs, path precompute)
    hipStream_t fix_stream = nullptr;
    hipEvent_t fix_fork = nullptr, fix_join = nullptr;
    uint16_t* d_fk[3] = {nullptr, nullptr, nullptr};
    uint64_t* d_fsorted[3] = {nullptr, nullptr, nullptr};
    uint8_t* d_ftmp[3] = {nullptr, nullptr, nullptr};
    size_t fk_cap[3] = {0, 0, 0}, fsorted_cap[3] = {0, 0, 0}, ftmp_cap[3] = {0, 0, 0};
    uint2* d_as_first = nullptr;     // first steps per branch from the counting walk
    size_t as_first_cap = 0;
    uint32_t* d_as_off32 = nullptr;  // 32-bit branch offsets and per-step cubes for the dirty test
    uint32_t* d_as_cube = nullptr;
    uint32_t* d_as_order = nullptr;  // branches in check order (first nb), sort input (second nb)
    uint64_t* d_as_sd = nullptr;     // per-step dirty bits
    uint32_t* d_as_fails = nullptr;  // per-branch bit: the last check found a mismatch
    size_t as_off32_cap = 0, as_cube_cap = 0, as_order_cap = 0, as_sd_cap = 0, as_fails_cap = 0;
    hipStream_t as_stream = nullptr;  // the ascending-path precompute, overlapping the first sweeps
    hipEvent_t as_ev_fork = nullptr, as_ev_count = nullptr, as_ev_done = nullptr;
    ull* h_as_total = nullptr;        // pinned: total steps of the precomputed paths, and (low
                                      // half of [1]) the branches without a step
    uint32_t* d_wall_scr = nullptr;  // wall walks: per-thread visited sets and stacks
    size_t wall_scr_cap = 0;
    uint32_t* d_wall_ovf = nullptr;  // wall walks to repeat with more slots (two lists)
    size_t wall_ovf_cap = 0;
    int64_t* d_labels = nullptr;  // min labels (2 per f-critical edge), max labels (2 per face)
    int64_t* d_walk = nullptr;    // per-branch walk results (4 per start cell)
    ull* d_ctr = nullptr;
    ull* h_ctr = nullptr;
    i64 cap_n = 0;   // vertices the buffers are sized for
    ull rec_cap = 0;
    ull fc_cap = 0;
    uint32_t epoch = 0;
    std::vector<int64_t> traj;
    double phase[6] = {0, 0, 0, 0, 0, 0};
    // live timing of the dense detection kernel on the context stream (HIP events)
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    double det_launches = 0, det_ms = 0, det_bytes = 0, det_voxels = 0;
    std::vector<dmtz_ctx*> workers;  // batch lanes: contexts of their own (stream, buffers)
    double dirty_total = 0;  // star codes re-evaluated incrementally in the last correction
    int64_t stalls = 0, finished = 0, radius = 0;  // finisher outcome of the last correction
    // precomputed ascending separatrix paths of f (geometry mode)
    ull* d_as_cnt = nullptr;   // comparison steps per branch, then scanned into d_as_off
    ull* d_as_off = nullptr;
    uint2* d_as_steps = nullptr;  // (tet, f partner face) per step
    uint32_t* d_cubes = nullptr;  // dirty-cube bitmap since the last path check
    uint32_t* d_vbits = nullptr;  // dirty-vertex bitmap since the last path check
    uint4* d_dsv = nullptr;       // per f-critical edge: endpoints and their f pairings
    uint32_t* d_dse = nullptr;    // per f-critical edge: its id
    int64_t* d_ds_res = nullptr;  // per f-critical edge: recorded vertex of the last check
    void* d_scan_tmp = nullptr;
    size_t as_cnt_cap = 0, as_off_cap = 0, as_steps_cap = 0, cubes_cap = 0, scan_tmp_cap = 0;
    size_t vbits_cap = 0, dsv_cap = 0, dse_cap = 0, ds_res_cap = 0;
};

namespace {

void free_bufs(dmtz_ctx* ctx) {
    (void)hipFree(ctx->d_f);
    (void)hipFree(ctx->d_g);
    (void)hipFree(ctx->d_counter);
    (void)hipFree(ctx->d_stamp);
    (void)hipFree(ctx->d_dstamp);
    (void)hipFree(ctx->d_targets);
    (void)hipFree(ctx->d_changed);
    (void)hipFree(ctx->d_dirty);
    (void)hipFree(ctx->d_code_f);
    (void)hipFree(ctx->d_code_g);
    (void)hipFree(ctx->d_critf);
    (void)hipFree(ctx->d_mm);
    (void)hipFree(ctx->d_tile_nz);
    ctx->d_tile_nz = nullptr;
    ctx->d_f = ctx->d_g = nullptr;
    ctx->d_counter = nullptr;
    ctx->d_stamp = ctx->d_dstamp = ctx->d_targets = ctx->d_changed = ctx->d_dirty = nullptr;
    ctx->d_code_f = ctx->d_code_g = nullptr;
    ctx->d_critf = ctx->d_mm = nullptr;
    ctx->field_bytes = 0;
    ctx->cap_n = 0;
}

void free_rec(dmtz_ctx* ctx) {
    (void)hipFree(ctx->d_rec);
    ctx->d_rec = nullptr;
    ctx->rec_cap = 0;
}

void free_paths(dmtz_ctx* ctx) {
    (void)hipFree(ctx->d_as_cnt);
    (void)hipFree(ctx->d_as_off);
    (void)hipFree(ctx->d_as_steps);
    (void)hipFree(ctx->d_cubes);
    (void)hipFree(ctx->d_vbits);
    (void)hipFree(ctx->d_dsv);
    (void)hipFree(ctx->d_dse);
    (void)hipFree(ctx->d_ds_res);
    (void)hipFree(ctx->d_scan_tmp);
    (void)hipFree(ctx->d_as_first);
    ctx->d_as_first = nullptr;
    ctx->as_first_cap = 0;
    (void)hipFree(ctx->d_as_off32);
    (void)hipFree(ctx->d_as_cube);
    (void)hipFree(ctx->d_as_order);
    (void)hipFree(ctx->d_as_sd);
    (void)hipFree(ctx->d_as_fails);
    ctx->d_as_off32 = ctx->d_as_cube = ctx->d_as_order = ctx->d_as_fails = nullptr;
    ctx->d_as_sd = nullptr;
    ctx->as_off32_cap = ctx->as_cube_cap = ctx->as_order_cap = ctx->as_sd_cap = ctx->as_fails_cap = 0;
    (void)hipFree(ctx->d_wall_scr);
    (void)hipFree(ctx->d_wall_ovf);
    ctx->d_wall_scr = ctx->d_wall_ovf = nullptr;
    ctx->wall_scr_cap = ctx->wall_ovf_cap = 0;
    ctx->d_as_cnt = ctx->d_as_off = nullptr;
    ctx->d_as_steps = nullptr;
    ctx->d_cubes = ctx->d_vbits = ctx->d_dse = nullptr;
    ctx->d_dsv = nullptr;
    ctx->d_ds_res = nullptr;
    ctx->d_scan_tmp = nullptr;
    ctx->as_cnt_cap = ctx->as_off_cap = ctx->as_steps_cap = ctx->cubes_cap = ctx->scan_tmp_cap = 0;
    ctx->vbits_cap = ctx->dsv_cap = ctx->dse_cap = ctx->ds_res_cap = 0;
}

void free_fc(dmtz_ctx* ctx) {
    (void)hipFree(ctx->d_fc);
    (void)hipFree(ctx->d_path);
    (void)hipFree(ctx->d_labels);
    (void)hipFree(ctx->d_walk);
    ctx->d_fc = nullptr;
    ctx->d_path = ctx->d_labels = ctx->d_walk = nullptr;
    ctx->fc_cap = 0;
}

#define CK(x)                                                                         \
    do {                                                                              \
        hipError_t e_ = (x);                                                          \
        if (e_ != hipSuccess) {                                                       \
            ctx->err = std::string(#x) + " failed: " + hipGetErrorString(e_);         \
            return DMTZ_E_HIP;                                                        \
        }                                                                             \
    } while (0)

int ensure(dmtz_ctx* ctx, i64 N, size_t elem, bool want_fields) {
    if (N > ctx->cap_n) {
        free_bufs(ctx);
        const size_t n = (size_t)N;
        CK(hipMalloc(&ctx->d_counter, n));
        CK(hipMalloc(&ctx->d_stamp, sizeof(uint32_t) * n));
        CK(hipMalloc(&ctx->d_dstamp, sizeof(uint32_t) * n));
        CK(hipMemsetAsync(ctx->d_stamp, 0, sizeof(uint32_t) * n, ctx->stream));
        CK(hipMemsetAsync(ctx->d_dstamp, 0, sizeof(uint32_t) * n, ctx->stream));
        CK(hipMalloc(&ctx->d_targets, sizeof(uint32_t) * n));
        CK(hipMalloc(&ctx->d_changed, sizeof(uint32_t) * n));
        CK(hipMalloc(&ctx->d_dirty, sizeof(uint32_t) * n));
        CK(hipMalloc(&ctx->d_code_f, sizeof(StarCode) * n));
        CK(hipMalloc(&ctx->d_code_g, sizeof(StarCode) * n));
        CK(hipMalloc(&ctx->d_critf, sizeof(CritMask) * n));
        CK(hipMalloc(&ctx->d_mm, sizeof(CritMask) * n));
        CK(hipMalloc(&ctx->d_tile_nz, sizeof(uint32_t) * (n / EMIT_TILE + 2)));
        ctx->cap_n = N;
        ctx->epoch = 0;
    }
    if (want_fields && ctx->field_bytes < (size_t)N * elem) {
        (void)hipFree(ctx->d_f);
        (void)hipFree(ctx->d_g);
        ctx->d_f = ctx->d_g = nullptr;
        CK(hipMalloc(&ctx->d_f, (size_t)N * elem));
        CK(hipMalloc(&ctx->d_g, (size_t)N * elem));
        ctx->field_bytes = (size_t)N * elem;
    }
    return 0;
}

// Record lists grow on demand: detection is recomputed from f and g, so an overflowing launch is
// simply repeated after the resize.
int grow_rec(dmtz_ctx* ctx, ull need) {
    if (need <= ctx->rec_cap) return 0;
    free_rec(ctx);
    ull cap = need + need / 4 + 4096;
    CK(hipMalloc(&ctx->d_rec, sizeof(uint64_t) * 4 * cap));
    ctx->rec_cap = cap;
    return 0;
}

int grow_fc(dmtz_ctx* ctx, ull need) {
    if (need <= ctx->fc_cap) return 0;
    free_fc(ctx);
    ull cap = need + need / 4 + 4096;
    CK(hipMalloc(&ctx->d_fc, sizeof(uint64_t) * 4 * cap));
    CK(hipMalloc(&ctx->d_path, sizeof(int64_t) * 6 * cap));
    CK(hipMalloc(&ctx->d_labels, sizeof(int64_t) * 4 * cap));
    CK(hipMalloc(&ctx->d_walk, sizeof(int64_t) * 4 * cap));
    ctx->fc_cap = cap;
    return 0;
}

// Grow a device buffer of P to at least need elements (by x1.25 + 4096).
template <class P>
int grow_buf(dmtz_ctx* ctx, P*& p, size_t& cap, size_t need) {
    if (need <= cap) return 0;
    (void)hipFree(p);
    p = nullptr;
    cap = 0;
    const size_t n = need + need / 4 + 4096;
    CK(hipMalloc(&p, sizeof(P) * n));
    cap = n;
    return 0;
}

int check_dims(dmtz_ctx* ctx, i64 nx, i64 ny, i64 nz) {
    if (!ctx) return DMTZ_E_ARG;
    if (nx < 3 || ny < 2 || nz < 2) {
        ctx->err = "grid must have nx >= 3, ny >= 2, nz >= 2";
        return DMTZ_E_ARG;
    }
    if (nx * ny * nz >= ((i64)1 << 31)) {
        ctx->err = "grid has 2^31 or more vertices";
        return DMTZ_E_ARG;
    }
    return 0;
}

Lists4 rec_lists(dmtz_ctx* ctx) {
    Lists4 L;
    for (int k = 0; k < 4; k++) L.l[k] = ctx->d_rec + (size_t)k * ctx->rec_cap;
    L.cap = ctx->rec_cap;
    return L;
}
Lists4 fcrit_lists(dmtz_ctx* ctx) {
    Lists4 L;
    for (int k = 0; k < 4; k++) L.l[k] = ctx->d_fc + (size_t)k * ctx->fc_cap;
    L.cap = ctx->fc_cap;
    return L;
}

ull max4(const ull* x) {
    ull m = x[0];
    for (int k = 1; k < 4; k++) m = x[k] > m ? x[k] : m;
    return m;
}

int sync_ctr(dmtz_ctx* ctx) {
    CK(hipMemcpyAsync(ctx->h_ctr, ctx->d_ctr, sizeof(ull) * K_NCTR, hipMemcpyDeviceToHost, ctx->stream));
    CK(hipStreamSynchronize(ctx->stream));
    return 0;
}

int overflow_check(dmtz_ctx* ctx) {
    for (int k = 0; k < 4; k++)
        if (ctx->h_ctr[K_MIN + k] > ctx->rec_cap || ctx->h_ctr[K_FC0 + k] > ctx->fc_cap) {
            ctx->err = "mismatch list capacity exceeded";
            return DMTZ_E_CAPACITY;
        }
    if (ctx->h_ctr[K_DS] > ctx->fc_cap || ctx->h_ctr[K_AS] > ctx->fc_cap || ctx->h_ctr[K_DSW] > ctx->fc_cap) {
        ctx->err = "separatrix list capacity exceeded";
        return DMTZ_E_CAPACITY;
    }
    if (ctx->h_ctr[K_TGT] > (ull)ctx->cap_n) {
        ctx->err = "edit set capacity exceeded";
        return DMTZ_E_CAPACITY;
    }
    return 0;
}

using Clock = std::chrono::steady_clock;
double secs(Clock::time_point a, Clock::time_point b) { return std::chrono::duration<double>(b - a).count(); }

template <class T>
struct Driver {
    dmtz_ctx* ctx;
    const T* f;
    T* g;
    Dims d;
    Counts c;
    double eps, range;
    dmtz_flags fl;

    // Detection of the four critical-cell mismatch kinds into the record lists.
    int detect() {
        if (ctx->rec_cap == 0) {
            int rc0 = grow_rec(ctx, (ull)(c.N / 4));
            if (rc0) return rc0;
        }
        for (;;) {
            int rc = detect_once();
            if (rc) return rc;
            ull m = max4(ctx->h_ctr + K_MIN);
            if (m <= ctx->rec_cap) break;
            rc = grow_rec(ctx, m);
            if (rc) return rc;
        }
        return overflow_check(ctx);
    }

    int detect_once() {
        k_zero_ctrs<<<1, 64, 0, ctx->stream>>>(ctx->d_ctr, ctr_bit(K_MIN) | ctr_bit(K_MIN + 1) | ctr_bit(K_MIN + 2) |
                                                        ctr_bit(K_MIN + 3) | ctr_bit(K_DTILES));
        CK(hipGetLastError());
        CK(hipEventRecord(ctx->ev0, ctx->stream));
        const ull ntiles = (ull)((c.N + EMIT_TILE - 1) / EMIT_TILE);
        k_detect<<<(unsigned)(ntiles < DETECT_BLOCKS ? ntiles : DETECT_BLOCKS), BLOCK, 0, ctx->stream>>>(
            ctx->d_mm, c.N, ctx->d_tile_nz, rec_lists(ctx), ctx->d_ctr + K_MIN, ctx->d_ctr + K_DTILES);
        CK(hipGetLastError());
        CK(hipEventRecord(ctx->ev1, ctx->stream));
        int rc = sync_ctr(ctx);
        if (rc) return rc;
        float ms = 0;
        CK(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
        // bytes the kernel must move: the tile counts, the masks of the tiles with a mismatch,
        // and one 8-byte record written per mismatch emitted
        ull m = ctx->h_ctr[0] + ctx->h_ctr[1] + ctx->h_ctr[2] + ctx->h_ctr[3];
        ctx->det_launches += 1;
        ctx->det_ms += ms;
        ctx->det_bytes += 4.0 * (double)ntiles + (double)ctx->h_ctr[K_DTILES] * (double)EMIT_TILE * sizeof(CritMask) +
                          8.0 * (double)m;
        ctx->det_voxels += (double)c.N;
        return 0;
    }

    TargetSink sink() {
        TargetSink s;
        s.stamp = ctx->d_stamp;
        s.epoch = ctx->epoch;
        s.list = ctx->d_targets;
        s.ctr = ctx->d_ctr + K_TGT;
        s.cap = (ull)ctx->cap_n;
        return s;
    }

    uint32_t zone_from = 0;  // first epoch of the unchanged sweeps since the last change (0: none)

    int begin_edits(bool for_apply = false) {
        ctx->epoch++;
        if (ctx->epoch == 0) {  // stamp wrap (after 2^32 edit sets): clear and restart; the finisher
                                // zone then keeps only the edit sets from here on
            CK(hipMemsetAsync(ctx->d_stamp, 0, sizeof(uint32_t) * (size_t)c.N, ctx->stream));
            CK(hipMemsetAsync(ctx->d_dstamp, 0, sizeof(uint32_t) * (size_t)c.N, ctx->stream));
            ctx->epoch = 1;
            if (zone_from) zone_from = 1;
        }
        // the edit set; with for_apply also the counters of its apply (changed and dirty lists,
        // moved edits), whose last values the host has read before this epoch begins (the
        // finisher keeps its changed list across its epochs, so it zeroes those itself)
        k_zero_ctrs<<<1, 64, 0, ctx->stream>>>(
            ctx->d_ctr, ctr_bit(K_TGT) | (for_apply ? ctr_bit(K_CHG) | ctr_bit(K_DIRTY) | ctr_bit(K_MOVED) : 0u));
        CK(hipGetLastError());
        return 0;
    }

    // Repair phase of one critical-point sweep (main_3d.cu:6801-6833) + apply (:6837). The
    // records of each kind are grouped by branch class, then the kinds are repaired concurrently
    // (the edits form a set: their order does not matter): minima and 2-saddles on the context's
    // stream, saddles then maxima on a second one, so that one kind's longest chains of dependent
    // reads overlap another's work. Two streams rather than one per kind: the process's streams
    // share four hardware queues, and two streams on one queue run one after the other.

    // Group the n records of kind ty (1..3) by branch class; recs then points at the grouped copy.
    // Grouping only orders the records, so lists beyond 32-bit slots are left as they are, and
    // so are short lists: a few blocks of records take as long as their slowest chain whatever
    // their order, and the grouping's five launches would cost more than it saves.
    static constexpr ull FIX_GROUP_MIN = 8192;
    int sort_class(int ty, const uint64_t*& recs, ull n) {
        if (n < FIX_GROUP_MIN || n >= (1ull << 32)) return 0;
        const int k = ty - 1;
        hipStream_t st = ctx->stream;
        int rc = grow_buf(ctx, ctx->d_fk[k], ctx->fk_cap[k], n);
        if (rc) return rc;
        rc = grow_buf(ctx, ctx->d_fsorted[k], ctx->fsorted_cap[k], n);
        if (rc) return rc;
        auto cls = ty == 1 ? k_fix_class<T, 1> : ty == 2 ? k_fix_class<T, 2> : k_fix_class<T, 3>;
        cls<<<nblocks(n), BLOCK, 0, st>>>(recs, n, f, ctx->d_code_f, g, ctx->d_code_g, ctx->d_critf, d, c,
                                          ctx->d_fk[k]);
        CK(hipGetLastError());
        const ull nblk = (n + CS_PER_BLOCK - 1) / CS_PER_BLOCK, m = nblk * CS_BINS;
        const size_t hist_bytes = (sizeof(uint32_t) * m + 255) / 256 * 256;
        size_t tmp = 0;
        CK(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp, (uint32_t*)nullptr, (uint32_t*)nullptr, (int)m, st));
        rc = grow_buf(ctx, ctx->d_ftmp[k], ctx->ftmp_cap[k], 2 * hist_bytes + tmp);
        if (rc) return rc;
        uint32_t* hist = reinterpret_cast<uint32_t*>(ctx->d_ftmp[k]);
        uint32_t* first = reinterpret_cast<uint32_t*>(ctx->d_ftmp[k] + hist_bytes);
        k_class_hist<<<(unsigned)nblk, BLOCK, 0, st>>>(ctx->d_fk[k], n, nblk, hist);
        CK(hipGetLastError());
        CK(hipcub::DeviceScan::ExclusiveSum(ctx->d_ftmp[k] + 2 * hist_bytes, tmp, hist, first, (int)m, st));
        k_class_scatter<<<(unsigned)nblk, BLOCK, 0, st>>>(ctx->d_fk[k], recs, n, nblk, first, ctx->d_fsorted[k]);
        CK(hipGetLastError());
        recs = ctx->d_fsorted[k];
        return 0;
    }

    int fix_cp(const ull counts[4]) {
        int rc = begin_edits(true);
        if (rc) return rc;
        ull nrec = 0;
        rc = repair_records(counts, nrec);
        if (rc) return rc;
        return apply(nrec);
    }

    // The repair rules over the current mismatch lists (counts[ty] records of type ty) into the
    // edit set of the current epoch; nrec receives the number of records repaired.
    int repair_records(const ull counts[4], ull& nrec) {
        Lists4 L = rec_lists(ctx);
        bool active[4];
        const uint64_t* recs[4];
        nrec = 0;
        int rc = 0;
        for (int ty = 0; ty < 4; ty++) {
            active[ty] = counts[ty] != 0 && !((ty == 1 || ty == 2) && fl.preserve_saddles != 1);
            recs[ty] = L.l[ty];
            if (!active[ty]) continue;
            nrec += counts[ty];
            if (ty != 0) {
                rc = sort_class(ty, recs[ty], counts[ty]);
                if (rc) return rc;
            }
        }
        const bool side = ctx->fix_concurrent && (active[1] || active[3]) && (active[0] || active[2]);
        if (side && !ctx->fix_stream) {
            // created on first use: batch lanes never run a second stream
            CK(hipStreamCreateWithFlags(&ctx->fix_stream, hipStreamNonBlocking));
            CK(hipEventCreateWithFlags(&ctx->fix_fork, hipEventDisableTiming));
            CK(hipEventCreateWithFlags(&ctx->fix_join, hipEventDisableTiming));
        }
        if (side) {
            CK(hipEventRecord(ctx->fix_fork, ctx->stream));
            CK(hipStreamWaitEvent(ctx->fix_stream, ctx->fix_fork, 0));
        }
        for (int ty : {2, 0, 1, 3}) {
            if (!active[ty]) continue;
            hipStream_t st = side && (ty == 1 || ty == 3) ? ctx->fix_stream : ctx->stream;
            auto kern = ty == 0 ? k_fix_cp<T, 0> : ty == 1 ? k_fix_cp<T, 1> : ty == 2 ? k_fix_cp<T, 2> : k_fix_cp<T, 3>;
            kern<<<nblocks(counts[ty]), BLOCK, 0, st>>>(recs[ty], counts[ty], f, ctx->d_code_f, g, ctx->d_code_g, d,
                                                        c, ctx->d_counter, eps, range, fl.edit_type, sink());
            CK(hipGetLastError());
        }
        if (side) {  // the context's stream waits for the second stream's kinds before the edits apply
            CK(hipEventRecord(ctx->fix_join, ctx->fix_stream));
            CK(hipStreamWaitEvent(ctx->stream, ctx->fix_join, 0));
        }
        return 0;
    }

    // applyDeltaBuffer1 (main_3d.cu:1122-1164) over the edit set, then re-evaluation of the g star
    // codes of the closed stars of the vertices whose value changed (the g gradient of
    // main_3d.cu:6864-6874 restricted to where it can change). Nothing here waits for the
    // device: the edit set's size stays on the device, nrec (the records repaired) sizes the
    // launches, and the count of edits that changed a value or a counter (K_MOVED) is read with
    // the next detection's counters. Each vertex enters the edit set at most once per epoch, so
    // the set never exceeds the N entries of its buffer.
    int apply(ull nrec) {  // its counters start at zero with the epoch (begin_edits(true))
        if (!nrec) return 0;  // no record, no edit
        const ull est = nrec < (ull)c.N ? nrec : (ull)c.N;
        k_apply<T><<<stride_blocks(est), BLOCK, 0, ctx->stream>>>(
            ctx->d_targets, ctx->d_ctr + K_TGT, (ull)ctx->cap_n, f, g, ctx->d_counter, eps, range, fl.preserve_type,
            fl.edit_type, ctx->d_changed, ctx->d_ctr + K_CHG, ctx->d_ctr + K_MOVED);
        CK(hipGetLastError());
        return refresh(est);
    }

    // Re-evaluate the g star codes of the closed stars of the vertices in the changed list (the g
    // gradient of main_3d.cu:6864-6874 restricted to where it can change). est, the host's
    // estimate of the list's length, sizes the launches; when it is large the exact length is
    // read back to choose between re-evaluating every star and the closed stars of the list
    // (both give the same codes).
    int refresh(ull est) {
        if (!est) return 0;
        if (16 * est >= (ull)c.N) {  // many stars may change: decide on the exact count
            int rc = sync_ctr(ctx);
            if (rc) return rc;
            est = ctx->h_ctr[K_CHG];
            if (!est) return 0;
        }
        if (16 * est >= (ull)c.N) {  // most stars change: re-evaluate all of them
            int rc = g_codes();
            if (rc) return rc;
            as_full = true;
            ctx->dirty_total += (double)c.N;
            return 0;
        }
        if (DIRTY_DENSE * 15 * est >= (ull)c.N) {  // long list: plain stores, then one pass over the stamps
            k_dirty_mark<<<stride_blocks(15 * est), BLOCK, 0, ctx->stream>>>(ctx->d_changed, ctx->d_ctr + K_CHG, d,
                                                                            ctx->d_dstamp, ctx->epoch);
            CK(hipGetLastError());
            const ull per_block = (ull)BLOCK * 4 * DIRTY_QUADS;
            k_dirty_collect<<<(unsigned)(((ull)c.N + per_block - 1) / per_block), BLOCK, 0, ctx->stream>>>(
                ctx->d_dstamp, c.N, ctx->epoch, ctx->d_dirty, ctx->d_ctr + K_DIRTY);
        } else {
            k_dirty<<<stride_blocks(15 * est), BLOCK, 0, ctx->stream>>>(ctx->d_changed, ctx->d_ctr + K_CHG, d,
                                                                       ctx->d_dstamp, ctx->epoch, ctx->d_dirty,
                                                                       ctx->d_ctr + K_DIRTY);
        }
        CK(hipGetLastError());
        k_code_list_g<T><<<stride_blocks(15 * est), BLOCK, 0, ctx->stream>>>(
            g, d, ctx->d_code_g, ctx->d_critf, ctx->d_mm, ctx->d_dirty, ctx->d_ctr + K_DIRTY, ctx->d_ctr + K_STARS,
            as_ready ? ctx->d_cubes : nullptr, as_ready ? ctx->d_vbits : nullptr, ctx->d_tile_nz);
        CK(hipGetLastError());
        return 0;
    }

    // Finisher step (no counterpart in the reference). The zone is the vertices of the first
    // cnt[k] mismatch records and the vertices targeted since zone_from. With r = ctx->radius
    // below the longest grid side and a non-empty zone: snap every vertex within Chebyshev
    // distance r of the zone that is not frozen yet, freeze them all, and repair the ties the
    // float snap creates by one-ulp raises. Otherwise restore g = f on every vertex and freeze it.
    int finish_step(const i64* cnt) {
        const int r = (int)ctx->radius;
        int rc = begin_edits();  // fresh epoch for the cell vertices; resets the list counter
        if (rc) return rc;
        const uint32_t from = zone_from, to = ctx->epoch;
        CK(hipMemsetAsync(ctx->d_ctr + K_CHG, 0, sizeof(ull) * 3, ctx->stream));  // changed, dirty, finished
        const ull ncell = (ull)(cnt[0] + cnt[1] + cnt[2] + cnt[3]);
        if (ncell) {
            k_zone_cells<<<nblocks(ncell), BLOCK, 0, ctx->stream>>>(rec_lists(ctx), (ull)cnt[0], (ull)cnt[1],
                                                                     (ull)cnt[2], (ull)cnt[3], d, ctx->d_stamp, to,
                                                                     ctx->d_targets, ctx->d_ctr + K_TGT);
            CK(hipGetLastError());
        }
        k_zone<<<nblocks(c.N), BLOCK, 0, ctx->stream>>>(ctx->d_stamp, from, to, c.N, ctx->d_targets,
                                                         ctx->d_ctr + K_TGT);
        CK(hipGetLastError());
        rc = sync_ctr(ctx);
        if (rc) return rc;
        const ull nz = ctx->h_ctr[K_TGT];
        const i64 span = d.W > d.H ? (d.W > d.D ? d.W : d.D) : (d.H > d.D ? d.H : d.D);
        if (nz == 0 || r >= span) {
            k_freeze_all<T><<<nblocks(c.N), BLOCK, 0, ctx->stream>>>(f, g, ctx->d_counter, c.N);
            CK(hipGetLastError());
            rc = g_codes();
            if (rc) return rc;
            as_full = true;
            ctx->finished += c.N;
        } else {
            rc = begin_edits();  // fresh epoch: each vertex is claimed once
            if (rc) return rc;
            const ull side = 2 * (ull)r + 1, total = nz * side * side * side;
            const ull blocks = nblocks(total) < 65536 ? nblocks(total) : 65536;
            k_snap_boxes<T><<<blocks, BLOCK, 0, ctx->stream>>>(ctx->d_targets, nz, r, d, f, g, ctx->d_counter, eps,
                                                                ctx->d_stamp, ctx->epoch, ctx->d_changed,
                                                                ctx->d_ctr + K_CHG);
            CK(hipGetLastError());
            rc = sync_ctr(ctx);
            if (rc) return rc;
            const ull nb = ctx->h_ctr[K_CHG];
            // Jacobi rounds of the tie repair; every round raises a vertex above a lower-f
            // neighbour, so at most nb rounds can be needed
            for (ull round = 0;; round++) {
                if (round > nb) {
                    ctx->err = "finisher tie repair did not settle";
                    return DMTZ_E_WALK;
                }
                rc = begin_edits();
                if (rc) return rc;
                CK(hipMemsetAsync(ctx->d_ctr + K_FINISH, 0, sizeof(ull), ctx->stream));
                k_tie_mark<T><<<nblocks(nb), BLOCK, 0, ctx->stream>>>(ctx->d_changed, nb, d, f, g, ctx->d_stamp,
                                                                       ctx->epoch, ctx->d_ctr + K_FINISH);
                CK(hipGetLastError());
                rc = sync_ctr(ctx);
                if (rc) return rc;
                if (ctx->h_ctr[K_FINISH] == 0) break;
                k_tie_raise<T><<<nblocks(nb), BLOCK, 0, ctx->stream>>>(ctx->d_changed, nb, g, ctx->d_stamp,
                                                                        ctx->epoch);
                CK(hipGetLastError());
            }
            rc = refresh(nb);
            if (rc) return rc;
            ctx->finished += (int64_t)nb;
        }
        zone_from = 0;
        return 0;
    }

    // Star codes of f and the f-critical status of every vertex's local cells (the gradient of
    // f, main_3d.cu:6539-6564, done once per run).
    int f_codes() {
        // the owned critical cells go to the mismatch masks first, which the g codes overwrite next
        int rc = codes(f, ctx->d_code_f, ctx->d_mm);
        if (rc) return rc;
        k_critf_gather<<<nblocks(c.N), BLOCK, 0, ctx->stream>>>(ctx->d_mm, d, c.N, ctx->d_critf);
        CK(hipGetLastError());
        return 0;
    }

    // g star codes and mismatch masks of every vertex (the g gradient, main_3d.cu:6624-6638);
    // needs the f-critical masks.
    int g_codes() {
        CK(hipMemsetAsync(ctx->d_tile_nz, 0, sizeof(uint32_t) * (size_t)((c.N + EMIT_TILE - 1) / EMIT_TILE),
                          ctx->stream));
        k_code_g<T><<<nblocks(c.N), BLOCK, 0, ctx->stream>>>(g, d, c, ctx->d_code_g, ctx->d_critf, ctx->d_mm,
                                                             ctx->d_tile_nz);
        CK(hipGetLastError());
        return 0;
    }

    // Star codes of every vertex of a field.
    int codes(const T* fld, StarCode* out, CritMask* own = nullptr) {
        k_code<T><<<nblocks(c.N), BLOCK, 0, ctx->stream>>>(fld, d, c, out, own);
        CK(hipGetLastError());
        return 0;
    }

    PathList path_list(int which) {
        PathList p;
        ull cap = ctx->fc_cap;
        p.id = ctx->d_path + (size_t)(2 * which) * cap;
        p.rec = ctx->d_path + (size_t)(2 * which + 1) * cap;
        p.ctr = ctx->d_ctr + (which == 0 ? K_DS : (which == 1 ? K_AS : K_DSW));
        p.cap = cap;
        return p;
    }

    ull nfce = 0, nfcf = 0;
    // Connectivity-mode ascending checks from g's critical faces instead of f's (dmtz_paths on
    // fields whose critical cells differ; in the correction loop the two sets are equal).
    const uint64_t* as_gfaces = nullptr;
    ull n_gfaces = 0;

    // f-critical cells of the types in type_mask into the f-critical lists (grown on demand).
    int fcrit(int type_mask) { return fcrit_of(ctx->d_code_f, type_mask); }

    // Critical cells of the types in type_mask of the field whose star codes are given, into the
    // f-critical lists (grown on demand).
    int fcrit_of(const StarCode* codes, int type_mask) {
        if (ctx->fc_cap == 0) {
            int rc0 = grow_fc(ctx, (ull)(c.N / 16));
            if (rc0) return rc0;
        }
        for (;;) {
            Lists4 FL = fcrit_lists(ctx);
            CK(hipMemsetAsync(ctx->d_ctr + K_FC0, 0, sizeof(ull) * 4, ctx->stream));
            k_fcrit<<<(unsigned)((c.N + EMIT_TILE - 1) / EMIT_TILE), BLOCK, 0, ctx->stream>>>(
                codes, c, FL, ctx->d_ctr + K_FC0, type_mask);
            CK(hipGetLastError());
            int rc = sync_ctr(ctx);
            if (rc) return rc;
            ull m = max4(ctx->h_ctr + K_FC0);
            if (m <= ctx->fc_cap) break;
            rc = grow_fc(ctx, m);
            if (rc) return rc;
        }
        return 0;
    }

    int prepare_paths() {
        int rc = fcrit(2 | 4);
        if (rc) return rc;
        Lists4 FL = fcrit_lists(ctx);
        nfce = ctx->h_ctr[K_FC0 + 1];
        nfcf = ctx->h_ctr[K_FC0 + 2];
        if (fl.preserve_vpath == 1 && fl.preserve_geometry == 0 && nfce + nfcf) {
            k_labels<T><<<nblocks(2 * (nfce + nfcf)), BLOCK, 0, ctx->stream>>>(
                FL.l[1], nfce, FL.l[2], nfcf, f, ctx->d_code_f, d, c, ctx->d_labels, ctx->d_labels + 2 * ctx->fc_cap,
                ctx->d_ctr + K_WALK);
            CK(hipGetLastError());
        }
        return 0;
    }

    // Separatrix detection (main_3d.cu:6955-6970): fills the ds / as lists.
    int detect_paths() {
        CK(hipMemsetAsync(ctx->d_ctr + K_DS, 0, sizeof(ull) * 2, ctx->stream));
        Lists4 FL = fcrit_lists(ctx);
        const bool conn = fl.preserve_geometry != 1;
        const bool pre = as_precomputable();
        if (pre && !as_ready && as_stage) {
            int rc = as_join();
            if (rc) return rc;
        }
        if (pre && !as_ready) {
            int rc = as_prepare();
            if (rc) return rc;
        }
        if (nfcf && pre) {
            int rc = as_check();
            if (rc) return rc;
        } else if (conn && as_gfaces) {
            if (n_gfaces) {
                k_as_branch<T><<<nblocks(2 * n_gfaces), BLOCK, 0, ctx->stream>>>(
                    as_gfaces, n_gfaces, nullptr, 1, f, ctx->d_code_f, g, ctx->d_code_g, d, c, ctx->d_walk,
                    ctx->d_ctr + K_WALK);
                CK(hipGetLastError());
            }
        } else if (nfcf) {
            const int64_t* lab = conn ? ctx->d_labels + 2 * ctx->fc_cap : nullptr;
            k_as_branch<T><<<nblocks(2 * nfcf), BLOCK, 0, ctx->stream>>>(FL.l[2], nfcf, lab, 0, f, ctx->d_code_f, g,
                                                                         ctx->d_code_g, d, c, ctx->d_walk,
                                                                         ctx->d_ctr + K_WALK);
            CK(hipGetLastError());
        }
        const ull nstarts = conn && as_gfaces ? n_gfaces : nfcf;
        if (nstarts && pre && as_compact) {
            k_as_combine_bits<<<nblocks((nstarts + 31) / 32), BLOCK, 0, ctx->stream>>>(ctx->d_walk, nstarts,
                                                                                      path_list(1), ctx->d_as_fails);
            CK(hipGetLastError());
        } else if (nstarts) {
            k_as_combine<<<nblocks(nstarts), BLOCK, 0, ctx->stream>>>(ctx->d_walk, nstarts, path_list(1), nullptr);
            CK(hipGetLastError());
        }
        if (nfce && pre) {
            int rc = ds_check();
            if (rc) return rc;
        } else if (nfce) {
            if (!conn) {
                k_ds_geo<T><<<nblocks(nfce), BLOCK, 0, ctx->stream>>>(FL.l[1], nfce, f, ctx->d_code_f, g, ctx->d_code_g,
                                                                      d, c, path_list(0));
                CK(hipGetLastError());
            } else {
                k_ds_branch<T><<<nblocks(2 * nfce), BLOCK, 0, ctx->stream>>>(FL.l[1], nfce, f, ctx->d_code_f, g,
                                                                             ctx->d_code_g, d, c, ctx->d_walk,
                                                                             ctx->d_ctr + K_WALK);
                CK(hipGetLastError());
                k_ds_combine<<<nblocks(nfce), BLOCK, 0, ctx->stream>>>(FL.l[1], nfce, ctx->d_walk, ctx->d_labels, d, c,
                                                                       path_list(0));
                CK(hipGetLastError());
            }
        }
        if (pre) {
            int rc = paths_checked();
            if (rc) return rc;
        }
        if (fl.preserve_connectors == 1) {
            int rc = walls();
            if (rc) return rc;
        }
        int rc = sync_ctr(ctx);
        if (rc) return rc;
        return overflow_check(ctx);
    }

    // 2-saddle walls (preserve_connectors, main_3d.cu:5064-5133) from every f-critical face
    // into the wall list. The walks run in tiers of visited-set size: every start with 256
    // slots, the walks that outgrow them with 8192, then with 2^18; a wall of more than 2^17
    // faces is a capacity error. Launches are chunked to keep the per-thread sets near 1 GiB.
    int walls() {
        CK(hipMemsetAsync(ctx->d_ctr + K_DSW, 0, sizeof(ull), ctx->stream));
        if (!nfcf) return 0;
        if (c.F >= (i64)0xFFFFFFFFll) {
            ctx->err = "2-saddle walls need fewer than 2^32 - 1 faces";
            return DMTZ_E_UNSUPPORTED;
        }
        static const uint32_t tiers[3] = {256u, 8192u, 1u << 18};
        const size_t budget = (size_t)1 << 28;  // uint32 entries of per-thread sets in flight
        int rc = grow_buf(ctx, ctx->d_wall_ovf, ctx->wall_ovf_cap, 2 * (size_t)nfcf);
        if (rc) return rc;
        Lists4 FL = fcrit_lists(ctx);
        const uint32_t* sel = nullptr;
        ull n = nfcf;
        for (int tier = 0; tier < 3 && n > 0; tier++) {
            const uint32_t slots = tiers[tier];
            const size_t per = 2 * (size_t)slots + 2;
            const ull chunk = (ull)(budget / per) < n ? (ull)(budget / per) : n;
            rc = grow_buf(ctx, ctx->d_wall_scr, ctx->wall_scr_cap, chunk * per);
            if (rc) return rc;
            uint32_t* ovf = ctx->d_wall_ovf + (tier % 2 == 0 ? 0 : nfcf);
            CK(hipMemsetAsync(ctx->d_ctr + K_WOVF, 0, sizeof(ull), ctx->stream));
            for (ull first = 0; first < n; first += chunk) {
                const ull last = first + chunk < n ? first + chunk : n;
                k_ds_wall<T><<<nblocks(last - first), BLOCK, 0, ctx->stream>>>(
                    FL.l[2], sel, first, last, f, ctx->d_code_f, g, ctx->d_code_g, d, c, ctx->d_wall_scr, slots,
                    path_list(2), ovf, ctx->d_ctr + K_WOVF);
                CK(hipGetLastError());
            }
            rc = sync_ctr(ctx);
            if (rc) return rc;
            n = ctx->h_ctr[K_WOVF];
            sel = ovf;
        }
        if (n > 0) {
            ctx->err = "a 2-saddle wall exceeds 2^17 faces";
            return DMTZ_E_CAPACITY;
        }
        return 0;
    }

    // Ascending geometry checks over precomputed f paths: in geometry mode, when tet and face
    // ids fit 32 bits.
    bool as_ready = false, as_full = true, as_compact = false;
    ull as_total = 0;  // steps of the compact path copies

    bool as_precomputable() const {
        return fl.preserve_geometry == 1 && c.T < ((i64)1 << 32) && c.F < ((i64)1 << 32) && c.E < ((i64)1 << 32) &&
               c.N < ((i64)1 << 32);
    }

    // Per-run precomputation of the separatrix checks (geometry mode): the ascending paths of f
    // (flat steps per branch), the descending per-edge data, and the dirty bitmaps.
    int as_prepare() {
        const ull nb = 2 * nfcf;
        Lists4 FL = fcrit_lists(ctx);
        int rc = nb ? as_paths(FL, nb) : 0;
        if (rc) return rc;
        rc = as_bitmaps_and_ds(FL, (size_t)(c.N >> 5) + 1, ctx->stream);
        if (rc) return rc;
        CK(hipStreamSynchronize(ctx->stream));
        return as_begin();
    }

    // The precomputed paths are in place: the checks start with a full one, and the branches
    // without a step, which the compact checks skip, keep an empty result.
    ull as_n0 = 0;  // branches without a step (the first as_n0 of the check order)
    int as_begin() {
        as_ready = true;
        as_full = true;
        as_n0 = as_compact ? (ctx->h_as_total[1] & 0xffffffffull) : 0;
        if (as_compact && nfcf) CK(hipMemsetAsync(ctx->d_walk, 0xFF, sizeof(int64_t) * 4 * nfcf, ctx->stream));
        const char* lg = std::getenv("DMTZ_LOG");
        if (lg && lg[0] == '1')
            std::fprintf(stderr, "dmtz ascending paths: %llu branches, %llu without a step, %llu steps\n",
                         (unsigned long long)(2 * nfcf), (unsigned long long)as_n0, (unsigned long long)as_total);
        return 0;
    }

    // The precompute depends on f alone, so single-field runs start it on a stream of its own
    // right after the first detection and let it overlap the critical-point sweeps; the first
    // separatrix check waits for it (as_join). Started after the first repair phase instead, it
    // slows the dense re-classification that follows (that kernel needs the wave slots the walk
    // holds) more than it spares the rules. as_stage: 0 none in flight, 1 counting walk enqueued,
    // 2 fill enqueued.
    int as_stage = 0;

    int as_start_async() {
        if (!ctx->fix_concurrent || !as_precomputable() || as_ready || !nfcf) return 0;
        if (!ctx->as_stream) {
            CK(hipStreamCreateWithFlags(&ctx->as_stream, hipStreamNonBlocking));
            CK(hipEventCreateWithFlags(&ctx->as_ev_fork, hipEventDisableTiming));
            CK(hipEventCreateWithFlags(&ctx->as_ev_count, hipEventDisableTiming));
            CK(hipEventCreateWithFlags(&ctx->as_ev_done, hipEventDisableTiming));
        }
        CK(hipEventRecord(ctx->as_ev_fork, ctx->stream));
        CK(hipStreamWaitEvent(ctx->as_stream, ctx->as_ev_fork, 0));
        int rc = as_paths_count(fcrit_lists(ctx), 2 * nfcf, ctx->as_stream, ctx->h_as_total);
        if (rc) return rc;
        CK(hipEventRecord(ctx->as_ev_count, ctx->as_stream));
        as_stage = 1;
        return 0;
    }

    // Enqueue the fill once the counting walk is done (wait = false: only if it already is).
    int as_advance(bool wait) {
        if (as_stage != 1) return 0;
        if (!wait) {
            const hipError_t q = hipEventQuery(ctx->as_ev_count);
            if (q == hipErrorNotReady) return 0;
            CK(q);
        } else {
            CK(hipEventSynchronize(ctx->as_ev_count));
        }
        Lists4 FL = fcrit_lists(ctx);
        int rc = as_paths_fill(FL, 2 * nfcf, ctx->as_stream, *ctx->h_as_total);
        if (rc) return rc;
        rc = as_bitmaps_and_ds(FL, (size_t)(c.N >> 5) + 1, ctx->as_stream);
        if (rc) return rc;
        CK(hipEventRecord(ctx->as_ev_done, ctx->as_stream));
        as_stage = 2;
        return 0;
    }

    // The context's stream waits for the precompute; the checks can then use it.
    int as_join() {
        int rc = as_advance(true);
        if (rc) return rc;
        if (as_stage == 2) {
            CK(hipEventSynchronize(ctx->as_ev_done));  // usually long done: the sweeps take longer
            CK(hipStreamWaitEvent(ctx->stream, ctx->as_ev_done, 0));
            as_stage = 0;
            return as_begin();
        }
        return 0;
    }

    // The ascending-path precompute in two phases on stream st: the counting walk and the scan of
    // the step counts (the total lands in *h_total when st gets there), then, with the total on
    // the host, the ordering, the fill and the compact copies.
    int as_paths(const Lists4& FL, ull nb) {
        int rc = as_paths_count(FL, nb, ctx->stream, ctx->h_as_total);
        if (rc) return rc;
        CK(hipStreamSynchronize(ctx->stream));
        return as_paths_fill(FL, nb, ctx->stream, *ctx->h_as_total);
    }

    int as_paths_count(const Lists4& FL, ull nb, hipStream_t st, ull* h_total) {
        int rc = grow_buf(ctx, ctx->d_as_cnt, ctx->as_cnt_cap, nb + 1);
        if (rc) return rc;
        rc = grow_buf(ctx, ctx->d_as_off, ctx->as_off_cap, nb + 1);
        if (rc) return rc;
        const int keep = as_keep_steps(nb);
        rc = grow_buf(ctx, ctx->d_as_first, ctx->as_first_cap, nb * (ull)keep);
        if (rc) return rc;
        CK(hipMemsetAsync(ctx->d_as_cnt + nb, 0, sizeof(ull), st));
        CK(hipMemsetAsync(ctx->d_ctr + K_ASNEXT, 0, sizeof(ull), st));
        const unsigned cap_blocks = st == ctx->stream ? AS_WS_BLOCKS : AS_WS_BLOCKS_SIDE;
        const unsigned wsb = nblocks(nb) < cap_blocks ? nblocks(nb) : cap_blocks;
        k_as_count_ws<T><<<wsb, BLOCK, 0, st>>>(FL.l[2], nb, f, ctx->d_code_f, d, c, ctx->d_as_cnt,
                                                          ctx->d_as_first, keep, ctx->d_ctr + K_ASNEXT,
                                                          ctx->d_ctr + K_WALK);
        CK(hipGetLastError());
        size_t tmp = 0;
        CK(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp, ctx->d_as_cnt, ctx->d_as_off, (int)(nb + 1), st));
        uint8_t* tp = (uint8_t*)ctx->d_scan_tmp;
        rc = grow_buf(ctx, tp, ctx->scan_tmp_cap, tmp);
        ctx->d_scan_tmp = tp;
        if (rc) return rc;
        CK(hipcub::DeviceScan::ExclusiveSum(ctx->d_scan_tmp, tmp, ctx->d_as_cnt, ctx->d_as_off, (int)(nb + 1),
                                            st));
        CK(hipMemcpyAsync(h_total, ctx->d_as_off + nb, sizeof(ull), hipMemcpyDeviceToHost, st));
        return 0;
    }

    int as_paths_fill(const Lists4& FL, ull nb, hipStream_t st, ull total) {
        const int keep = as_keep_steps(nb);
        size_t tmp = 0;
        CK(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp, ctx->d_as_cnt, ctx->d_as_off, (int)(nb + 1), st));
        int rc = grow_buf(ctx, ctx->d_as_steps, ctx->as_steps_cap, total);
        if (rc) return rc;
        as_compact = total < ((ull)1 << 32) && nb < ((ull)1 << 31);
        if (as_compact) {  // check order: by region, then path length (for the fill and the checks)
            rc = grow_buf(ctx, ctx->d_as_order, ctx->as_order_cap, nb);
            if (rc) return rc;
            const ull nblk = (nb + CS_PER_BLOCK - 1) / CS_PER_BLOCK, m = nblk * LEN_BINS;
            const size_t hist_bytes = (sizeof(uint32_t) * m + 255) / 256 * 256;
            size_t tmp2 = 0;
            CK(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp2, (uint32_t*)nullptr, (uint32_t*)nullptr, (int)m, st));
            uint8_t* tp2 = (uint8_t*)ctx->d_scan_tmp;
            rc = grow_buf(ctx, tp2, ctx->scan_tmp_cap, 2 * hist_bytes + (tmp > tmp2 ? tmp : tmp2));
            ctx->d_scan_tmp = tp2;
            if (rc) return rc;
            uint32_t* hist = reinterpret_cast<uint32_t*>(tp2);
            uint32_t* first = reinterpret_cast<uint32_t*>(tp2 + hist_bytes);
            k_len_hist<<<(unsigned)nblk, BLOCK, 0, st>>>(nb, ctx->d_as_cnt, nblk, hist);
            CK(hipGetLastError());
            CK(hipcub::DeviceScan::ExclusiveSum(tp2 + 2 * hist_bytes, tmp2, hist, first, (int)m, st));
            k_len_scatter<<<(unsigned)nblk, BLOCK, 0, st>>>(nb, ctx->d_as_cnt, nblk, first, ctx->d_as_order);
            CK(hipGetLastError());
            // branches without a step: the first slot of length 1 (low half of h_as_total[1])
            ctx->h_as_total[1] = 0;
            CK(hipMemcpyAsync(ctx->h_as_total + 1, first + nblk, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
            // steps in check order: offsets from the counts taken in that order
            k_as_gather_cnt<<<nblocks(nb + 1), BLOCK, 0, st>>>(nb, ctx->d_as_order, ctx->d_as_cnt,
                                                                       ctx->d_as_off);
            CK(hipGetLastError());
            CK(hipcub::DeviceScan::ExclusiveSum(tp2 + 2 * hist_bytes, tmp, ctx->d_as_off, ctx->d_as_cnt,
                                                (int)(nb + 1), st));
            std::swap(ctx->d_as_cnt, ctx->d_as_off);
            std::swap(ctx->as_cnt_cap, ctx->as_off_cap);
        }
        k_as_fill<T><<<nblocks(nb), BLOCK, 0, st>>>(FL.l[2], nb, as_compact ? ctx->d_as_order : nullptr, f,
                                                              ctx->d_code_f, d, c, ctx->d_as_off, ctx->d_as_first,
                                                              keep, ctx->d_as_steps);
        CK(hipGetLastError());
        if (as_compact) {
            rc = grow_buf(ctx, ctx->d_as_off32, ctx->as_off32_cap, nb + 1);
            if (rc) return rc;
            rc = grow_buf(ctx, ctx->d_as_cube, ctx->as_cube_cap, total + 1);
            if (rc) return rc;
            k_as_compact<<<nblocks(nb), BLOCK, 0, st>>>(nb, ctx->d_as_off, ctx->d_as_steps, ctx->d_as_off32,
                                                                  ctx->d_as_cube);
            CK(hipGetLastError());
            rc = grow_buf(ctx, ctx->d_as_sd, ctx->as_sd_cap, (total + 63) / 64 + 1);
            if (rc) return rc;
            rc = grow_buf(ctx, ctx->d_as_fails, ctx->as_fails_cap, (nb + 31) / 32 + 1);
            if (rc) return rc;
            as_total = total;
        }
        return 0;
    }

    int as_bitmaps_and_ds(const Lists4& FL, size_t words, hipStream_t st) {
        int rc = grow_buf(ctx, ctx->d_cubes, ctx->cubes_cap, words);
        if (rc) return rc;
        rc = grow_buf(ctx, ctx->d_vbits, ctx->vbits_cap, words);
        if (rc) return rc;
        CK(hipMemsetAsync(ctx->d_cubes, 0, sizeof(uint32_t) * words, st));
        CK(hipMemsetAsync(ctx->d_vbits, 0, sizeof(uint32_t) * words, st));
        if (nfce) {
            rc = grow_buf(ctx, ctx->d_dsv, ctx->dsv_cap, nfce);
            if (rc) return rc;
            rc = grow_buf(ctx, ctx->d_dse, ctx->dse_cap, nfce);
            if (rc) return rc;
            rc = grow_buf(ctx, ctx->d_ds_res, ctx->ds_res_cap, nfce);
            if (rc) return rc;
            k_ds_prepare<T><<<nblocks(nfce), BLOCK, 0, st>>>(FL.l[1], nfce, f, ctx->d_code_f, d, c,
                                                                       ctx->d_dsv, ctx->d_dse);
            CK(hipGetLastError());
        }
        return 0;
    }

    int as_check() {
        if (!as_ready) {
            int rc = as_prepare();
            if (rc) return rc;
        }
        const ull nb = 2 * nfcf;
        if (!nb) return 0;
        if (as_compact) {
            if (as_full) CK(hipMemsetAsync(ctx->d_as_fails, 0, sizeof(uint32_t) * ((nb + 31) / 32), ctx->stream));
            if (!as_full && as_total) {
                k_as_step_dirty<<<nblocks((as_total + SD_WORDS_PER_WAVE - 1) / SD_WORDS_PER_WAVE), BLOCK, 0,
                                  ctx->stream>>>(as_total, ctx->d_as_cube, ctx->d_cubes, ctx->d_as_sd);
                CK(hipGetLastError());
            }
            if (nb > as_n0)
                k_as_check32<T><<<nblocks(nb - as_n0), BLOCK, 0, ctx->stream>>>(
                    as_n0, nb, ctx->d_as_order, ctx->d_as_off32, ctx->d_as_sd, ctx->d_as_steps, g, ctx->d_code_g, d, c,
                    as_full ? 1 : 0, ctx->d_walk, ctx->d_as_fails);
        } else
            k_as_check<T><<<nblocks(nb), BLOCK, 0, ctx->stream>>>(nb, ctx->d_as_off, ctx->d_as_steps, g,
                                                                   ctx->d_code_g, d, c, ctx->d_cubes, as_full ? 1 : 0,
                                                                   ctx->d_walk);
        CK(hipGetLastError());
        return 0;
    }

    int ds_check() {
        k_ds_check<T><<<nblocks(nfce), BLOCK, 0, ctx->stream>>>(nfce, ctx->d_dsv, ctx->d_dse, g, ctx->d_code_g, d, c,
                                                                 ctx->d_vbits, as_full ? 1 : 0, ctx->d_ds_res,
                                                                 path_list(0));
        CK(hipGetLastError());
        return 0;
    }

    // After a path check: both dirty bitmaps start empty and the next check is incremental.
    int paths_checked() {
        const size_t words = (size_t)(c.N >> 5) + 1;
        CK(hipMemsetAsync(ctx->d_cubes, 0, sizeof(uint32_t) * words, ctx->stream));
        CK(hipMemsetAsync(ctx->d_vbits, 0, sizeof(uint32_t) * words, ctx->stream));
        as_full = false;
        return 0;
    }

    int fix_paths(ull nds, ull nas, ull nw) {
        int rc = begin_edits(true);
        if (rc) return rc;
        if (fl.preserve_connectors != 1) nw = 0;
        if (nds + nas + nw) {
            k_fix_paths<T><<<nblocks(nds + nas + nw), BLOCK, 0, ctx->stream>>>(
                path_list(1).rec, nas, path_list(0).rec, nds, path_list(2).rec, nw, f, ctx->d_code_f, g,
                ctx->d_code_g, d, c, ctx->d_counter, eps, range, fl.edit_type, sink());
            CK(hipGetLastError());
        }
        return apply(nds + nas + nw);
    }

    int filtration() {
        int rc = fcrit(15);
        if (rc) return rc;
        Lists4 FL = fcrit_lists(ctx);
        for (int ty = 0; ty < 4; ty++) {
            ull n = ctx->h_ctr[K_FC0 + ty];
            if (!n) continue;
            k_pin<T><<<nblocks(n), BLOCK, 0, ctx->stream>>>(FL.l[ty], n, f, g, d, c, eps);
            CK(hipGetLastError());
        }
        return 0;
    }

    // The loop of main_3d.cu:6521-7149, with an exit at a stall: a sweep is a function of
    // (g, counter, counts, sweep kind), so a sweep that changes no value or counter and whose
    // (counts, kind) repeats one seen since the last change starts a cycle the reference never
    // leaves. Without the finisher the run returns DMTZ_STALLED there; with it, finish_step() is
    // applied and the sweep's counts are re-evaluated.
    int run(dmtz_counts* out, int32_t* sweeps_out) {
        if (ctx->as_stream) CK(hipStreamSynchronize(ctx->as_stream));  // after a failed earlier run
        ctx->traj.clear();
        for (double& p : ctx->phase) p = 0;
        ctx->stalls = ctx->finished = ctx->radius = 0;
        zone_from = 0;
        auto t_start = Clock::now();
        int maxs = fl.max_sweeps > 0 ? fl.max_sweeps : 100000;
        CK(hipMemsetAsync(ctx->d_ctr, 0, sizeof(ull) * K_NCTR, ctx->stream));
        ctx->dirty_total = 0;
        auto t0 = Clock::now();
        int rc = f_codes();  // gradient of f, main_3d.cu:6539-6564
        if (rc) return rc;
        if (fl.filtration == 1) {
            rc = filtration();
            if (rc) return rc;
        }
        CK(hipMemsetAsync(ctx->d_counter, 1, (size_t)c.N, ctx->stream));  // init_counter :6782
        bool paths = fl.preserve_vpath == 1;
        if (paths) {
            rc = prepare_paths();
            if (rc) return rc;
        }
        rc = g_codes();  // gradient of g, main_3d.cu:6624-6638
        if (rc) return rc;
        CK(hipStreamSynchronize(ctx->stream));
        auto t1 = Clock::now();
        ctx->phase[0] += secs(t0, t1);
        rc = detect();
        if (rc) return rc;
        ctx->phase[1] += secs(t1, Clock::now());
        i64 cnt[8] = {(i64)ctx->h_ctr[0], (i64)ctx->h_ctr[1], (i64)ctx->h_ctr[2], (i64)ctx->h_ctr[3], 1, 1, 1, 0};
        if (paths) {
            rc = as_start_async();
            if (rc) return rc;
        }
        i64 sweeps = 0;
        int status = 0;
        const char* lg = std::getenv("DMTZ_LOG");
        const bool log = lg && lg[0] == '1';
        auto record = [&]() {
            for (int k = 0; k < 8; k++) ctx->traj.push_back(cnt[k]);
            if (log)
                std::fprintf(stderr, "dmtz sweep %lld: %lld %lld %lld %lld | %lld %lld  (%.3f s, %lld edits, %.0f stars)\n",
                             (long long)sweeps, (long long)cnt[0], (long long)cnt[1], (long long)cnt[2],
                             (long long)cnt[3], (long long)cnt[4], (long long)cnt[5], secs(t_start, Clock::now()),
                             (long long)ctx->h_ctr[K_TGT], ctx->dirty_total + (double)ctx->h_ctr[K_STARS]);
        };
        // Re-evaluation after the apply of a critical-point sweep (main_3d.cu:6864-6921).
        auto eval_cp = [&]() -> int {
            auto tb = Clock::now();
            int r = detect();
            if (r) return r;
            ctx->phase[1] += secs(tb, Clock::now());
            cnt[0] = (i64)ctx->h_ctr[0];
            cnt[1] = fl.preserve_saddles == 1 ? (i64)ctx->h_ctr[1] : 0;
            cnt[2] = fl.preserve_saddles == 1 ? (i64)ctx->h_ctr[2] : 0;
            cnt[3] = (i64)ctx->h_ctr[3];
            return 0;
        };
        // Re-evaluation after the apply of a separatrix sweep (main_3d.cu:7069-7143).
        auto eval_path = [&]() -> int {
            auto tc = Clock::now();
            int r = detect();
            if (r) return r;
            auto td = Clock::now();
            ctx->phase[1] += secs(tc, td);
            cnt[0] = (i64)ctx->h_ctr[0];
            cnt[1] = (i64)ctx->h_ctr[1];
            cnt[2] = (i64)ctx->h_ctr[2];
            cnt[3] = (i64)ctx->h_ctr[3];
            cnt[4] = cnt[5] = cnt[6] = cnt[7] = 0;
            if (cnt[0] == 0 && cnt[1] == 0 && cnt[2] == 0 && cnt[3] == 0) {
                r = detect_paths();
                if (r) return r;
                cnt[4] = (i64)ctx->h_ctr[K_DS];
                cnt[5] = (i64)ctx->h_ctr[K_AS];
                cnt[6] = fl.preserve_connectors == 1 ? (i64)ctx->h_ctr[K_DSW] : 0;
                ctx->phase[2] += secs(td, Clock::now());
            }
            return 0;
        };
        std::vector<std::vector<i64>> seen;  // (counts, kind) since the last change
        // After the re-evaluation of a sweep of the given kind (0 critical points, 1 separatrices).
        // Sets stalled when the run stalls without the finisher.
        bool stalled = false;
        auto after_sweep = [&](int kind) -> int {
            std::vector<i64> key(cnt, cnt + 8);
            key.push_back(kind);
            if (ctx->h_ctr[K_MOVED] > 0) {  // edits of the sweep that changed a value or a counter
                seen.assign(1, key);
                zone_from = 0;
                return 0;
            }
            if (zone_from == 0) zone_from = ctx->epoch;  // this sweep's edit set opens the zone
            bool any = false;
            for (int k = 0; k < 8; k++) any = any || cnt[k] != 0;
            if (!any) return 0;
            for (const auto& sk : seen)
                if (sk == key) {
                    ctx->stalls++;
                    if (fl.finish != 1) {
                        stalled = true;
                        return 0;
                    }
                    ctx->radius = ctx->radius ? 2 * ctx->radius : 1;  // 1, 2, 4, ...
                    int r = finish_step(cnt);
                    if (r) return r;
                    r = kind == 0 ? eval_cp() : eval_path();
                    if (r) return r;
                    key.assign(cnt, cnt + 8);
                    key.push_back(kind);
                    seen.assign(1, key);
                    return 0;
                }
            seen.push_back(key);
            return 0;
        };
        while (cnt[5] > 0 || cnt[4] > 0 || cnt[6] > 0 || cnt[2] > 0 || cnt[3] > 0 || cnt[0] > 0 || cnt[1] > 0) {
            while (cnt[3] > 0 || cnt[2] > 0 || cnt[0] > 0 || cnt[1] > 0) {
                if (sweeps >= maxs) {
                    status = DMTZ_NOT_CONVERGED;
                    goto done;
                }
                sweeps++;
                auto ta = Clock::now();
                ull cc[4] = {(ull)cnt[0], (ull)cnt[1], (ull)cnt[2], (ull)cnt[3]};
                rc = fix_cp(cc);
                if (rc) return rc;
                ctx->phase[3] += secs(ta, Clock::now());
                rc = eval_cp();
                if (rc) return rc;
                rc = after_sweep(0);
                if (rc) return rc;
                record();
                if (stalled) {
                    status = DMTZ_STALLED;
                    goto done;
                }
                rc = as_advance(false);
                if (rc) return rc;
            }
            if (!paths) break;
            if (sweeps >= maxs) {
                status = DMTZ_NOT_CONVERGED;
                goto done;
            }
            sweeps++;
            {
                auto ta = Clock::now();
                rc = detect_paths();
                if (rc) return rc;
                auto tb = Clock::now();
                ctx->phase[2] += secs(ta, tb);
                ull nds = ctx->h_ctr[K_DS], nas = ctx->h_ctr[K_AS], nw = ctx->h_ctr[K_DSW];
                rc = fix_paths(nds, nas, nw);
                if (rc) return rc;
                ctx->phase[4] += secs(tb, Clock::now());
                rc = eval_path();
                if (rc) return rc;
                rc = after_sweep(1);
                if (rc) return rc;
                record();
                if (stalled) {
                    status = DMTZ_STALLED;
                    goto done;
                }
            }
        }
    done:
        if (!paths) cnt[4] = cnt[5] = cnt[6] = cnt[7] = 0;
        if (as_stage) {  // a precompute the run did not need: let it finish, it is not used
            CK(hipStreamSynchronize(ctx->as_stream));
            as_stage = 0;
        }
        CK(hipStreamSynchronize(ctx->stream));
        ctx->phase[5] = secs(t_start, Clock::now());
        if (out) {
            out->min = cnt[0];
            out->saddle1 = cnt[1];
            out->saddle2 = cnt[2];
            out->max = cnt[3];
            out->ds_vpath = cnt[4];
            out->as_vpath = cnt[5];
            out->ds_wall = cnt[6];
            out->as_wall = cnt[7];
        }
        if (sweeps_out) *sweeps_out = (int32_t)sweeps;
        CK(hipMemcpyAsync(ctx->h_ctr, ctx->d_ctr, sizeof(ull) * K_NCTR, hipMemcpyDeviceToHost, ctx->stream));
        CK(hipStreamSynchronize(ctx->stream));
        ctx->dirty_total += (double)ctx->h_ctr[K_STARS];
        if (ctx->h_ctr[K_WALK] > 0) {
            ctx->err = "a separatrix walk exceeded N steps (cyclic gradient path)";
            return DMTZ_E_WALK;
        }
        return status;
    }
};

template <class T>
int correct_dev(dmtz_ctx* ctx, const T* d_f, T* d_g, i64 nx, i64 ny, i64 nz, double eps, double range,
                const dmtz_flags* flags, dmtz_counts* out, uint8_t* d_counter_out, int32_t* sweeps_out) {
    int rc = check_dims(ctx, nx, ny, nz);
    if (rc) return rc;
    if (!flags || !d_f || !d_g) {
        ctx->err = "null argument";
        return DMTZ_E_ARG;
    }
    if (flags->mode != 0) {
        ctx->err = "unknown mode";
        return DMTZ_E_ARG;
    }
    Dims d{nx, ny, nz};
    Counts c = make_counts(d);
    rc = ensure(ctx, c.N, sizeof(T), false);
    if (rc) return rc;
    Driver<T> drv{ctx, d_f, d_g, d, c, eps, range, *flags};
    rc = drv.run(out, sweeps_out);
    if (rc < 0) return rc;
    if (d_counter_out) {
        hipError_t e = hipMemcpyAsync(d_counter_out, ctx->d_counter, (size_t)c.N, hipMemcpyDeviceToDevice, ctx->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) {
            ctx->err = hipGetErrorString(e);
            return DMTZ_E_HIP;
        }
    }
    return rc;
}

template <class T>
int correct_host(dmtz_ctx* ctx, const T* f, T* g, i64 nx, i64 ny, i64 nz, double eps, double range,
                 const dmtz_flags* flags, dmtz_counts* out, uint8_t* counter_out, int32_t* sweeps_out) {
    int rc = check_dims(ctx, nx, ny, nz);
    if (rc) return rc;
    if (!f || !g) {
        ctx->err = "null argument";
        return DMTZ_E_ARG;
    }
    i64 N = nx * ny * nz;
    rc = ensure(ctx, N, sizeof(T), true);
    if (rc) return rc;
    CK(hipMemcpyAsync(ctx->d_f, f, sizeof(T) * (size_t)N, hipMemcpyHostToDevice, ctx->stream));
    CK(hipMemcpyAsync(ctx->d_g, g, sizeof(T) * (size_t)N, hipMemcpyHostToDevice, ctx->stream));
    int st = correct_dev<T>(ctx, (const T*)ctx->d_f, (T*)ctx->d_g, nx, ny, nz, eps, range, flags, out, nullptr,
                            sweeps_out);
    if (st < 0) return st;
    CK(hipMemcpyAsync(g, ctx->d_g, sizeof(T) * (size_t)N, hipMemcpyDeviceToHost, ctx->stream));
    if (counter_out)
        CK(hipMemcpyAsync(counter_out, ctx->d_counter, (size_t)N, hipMemcpyDeviceToHost, ctx->stream));
    CK(hipStreamSynchronize(ctx->stream));
    return st;
}

// The classifier alone (dmtz_classify_*): star codes of the field, then counts and masks of the
// cells left unpaired.
template <class T>
int classify_host(dmtz_ctx* ctx, const T* fld, i64 nx, i64 ny, i64 nz, int64_t* counts4, uint64_t* crit_lo,
                  uint32_t* crit_hi) {
    int rc = check_dims(ctx, nx, ny, nz);
    if (rc) return rc;
    if (!fld || !counts4) {
        ctx->err = "null argument";
        return DMTZ_E_ARG;
    }
    const i64 N = nx * ny * nz;
    rc = ensure(ctx, N, sizeof(T), true);
    if (rc) return rc;
    Dims d{nx, ny, nz};
    Counts c = make_counts(d);
    CK(hipMemcpyAsync(ctx->d_f, fld, sizeof(T) * (size_t)N, hipMemcpyHostToDevice, ctx->stream));
    k_code<T><<<nblocks(N), BLOCK, 0, ctx->stream>>>((const T*)ctx->d_f, d, c, ctx->d_code_f, nullptr);
    CK(hipGetLastError());
    CK(hipMemsetAsync(ctx->d_ctr, 0, sizeof(ull) * 4, ctx->stream));
    k_crit_count<<<nblocks(N), BLOCK, 0, ctx->stream>>>(ctx->d_code_f, N, ctx->d_ctr);
    CK(hipGetLastError());
    rc = sync_ctr(ctx);
    if (rc) return rc;
    for (int k = 0; k < 4; k++) counts4[k] = (int64_t)ctx->h_ctr[k];
    if (crit_lo || crit_hi) {
        k_critf_own<<<nblocks(N), BLOCK, 0, ctx->stream>>>(ctx->d_code_f, N, ctx->d_critf);
        CK(hipGetLastError());
        std::vector<CritMask> hm((size_t)N);
        CK(hipMemcpyAsync(hm.data(), ctx->d_critf, sizeof(CritMask) * (size_t)N, hipMemcpyDeviceToHost,
                          ctx->stream));
        CK(hipStreamSynchronize(ctx->stream));
        for (i64 v = 0; v < N; v++) {
            if (crit_lo) crit_lo[v] = hm[(size_t)v].lo;
            if (crit_hi) crit_hi[v] = hm[(size_t)v].hi;
        }
    }
    return 0;
}

// Batch correction (dmtz_correct_batch_*). Fields run concurrently in up to BATCH_LANES lanes,
// each a host thread with a worker context of its own (its own stream and buffers), so the
// sweeps of different fields overlap on the device and their per-sweep host round trips
// overlap too. Lanes take the fields largest first from a shared counter. Every field's result
// is what a single correction gives (each run is deterministic and lanes share nothing).
constexpr int BATCH_LANES = 8;

template <class T, class Field>
int correct_one(dmtz_ctx* w, const Field& fd, const dmtz_flags* flags, dmtz_counts* cnt, int32_t* sw) {
    return fd.on_device ? correct_dev<T>(w, fd.f, fd.g, fd.nx, fd.ny, fd.nz, fd.eps, fd.range, flags, cnt, nullptr, sw)
                        : correct_host<T>(w, fd.f, fd.g, fd.nx, fd.ny, fd.nz, fd.eps, fd.range, flags, cnt, nullptr,
                                          sw);
}

template <class T, class Field>
int correct_batch(dmtz_ctx* ctx, const Field* fields, int32_t n, const dmtz_flags* flags, dmtz_counts* counts_out,
                  int32_t* status_out, int32_t* sweeps_out) {
    if (!ctx || !flags || (n > 0 && !fields) || n < 0) {
        if (ctx) ctx->err = "null or invalid argument";
        return DMTZ_E_ARG;
    }
    if (n == 0) return 0;
    const int lanes = n < BATCH_LANES ? n : BATCH_LANES;
    while ((int)ctx->workers.size() < lanes) {
        dmtz_ctx* w = nullptr;
        const int rc = dmtz_ctx_create(&w, ctx->device);
        if (rc) {
            ctx->err = "batch lane context creation failed";
            return rc;
        }
        w->fix_concurrent = false;  // one stream per lane: more would share the four hardware queues
        ctx->workers.push_back(w);
    }
    for (dmtz_ctx* w : ctx->workers) w->err.clear();
    std::vector<int32_t> order(n);
    for (int32_t i = 0; i < n; i++) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int32_t a, int32_t b) {
        return fields[a].nx * fields[a].ny * fields[a].nz > fields[b].nx * fields[b].ny * fields[b].nz;
    });
    std::vector<int> rcs(n, DMTZ_E_NOT_RUN);  // fields that never start keep this status
    std::vector<dmtz_counts> cnts(n);
    std::vector<int32_t> sws(n, 0);
    std::atomic<int32_t> next{0};
    std::atomic<bool> failed{false};
    auto lane = [&](dmtz_ctx* w) {
        (void)hipSetDevice(w->device);
        for (;;) {
            const int32_t k = next.fetch_add(1);
            if (k >= n || failed.load()) break;
            const int32_t i = order[k];
            rcs[i] = correct_one<T>(w, fields[i], flags, &cnts[i], &sws[i]);
            if (rcs[i] < 0) failed.store(true);
        }
    };
    std::vector<std::thread> threads;
    for (int l = 1; l < lanes; l++) threads.emplace_back(lane, ctx->workers[l]);
    lane(ctx->workers[0]);
    for (auto& t : threads) t.join();
    (void)hipSetDevice(ctx->device);
    int worst = 0;
    for (int l = 0; l < lanes; l++) {  // kernel statistics of the lanes go to the caller's context
        dmtz_ctx* w = ctx->workers[l];
        ctx->det_launches += w->det_launches;
        ctx->det_ms += w->det_ms;
        ctx->det_bytes += w->det_bytes;
        ctx->det_voxels += w->det_voxels;
        w->det_launches = w->det_ms = w->det_bytes = w->det_voxels = 0;
    }
    int first_err = 0;  // first error in field order (DMTZ_E_NOT_RUN only follows an error)
    for (int32_t i = 0; i < n; i++) {
        if (status_out) status_out[i] = rcs[i];
        if (rcs[i] < 0) {
            if (rcs[i] != DMTZ_E_NOT_RUN && first_err == 0) first_err = rcs[i];
            continue;
        }
        if (counts_out) counts_out[i] = cnts[i];
        if (sweeps_out) sweeps_out[i] = sws[i];
        if (rcs[i] != 0) worst = 1;
    }
    if (first_err) {
        for (dmtz_ctx* w : ctx->workers)
            if (!w->err.empty()) ctx->err = w->err;
        return first_err;
    }
    return worst;
}

template <class T>
int gradient_host(dmtz_ctx* ctx, const T* fld, i64 nx, i64 ny, i64 nz, int64_t* pv, int64_t* pe, int64_t* pf,
                  int64_t* pt) {
    int rc = check_dims(ctx, nx, ny, nz);
    if (rc) return rc;
    Dims d{nx, ny, nz};
    Counts c = make_counts(d);
    rc = ensure(ctx, c.N, sizeof(T), true);
    if (rc) return rc;
    CK(hipMemcpyAsync(ctx->d_f, fld, sizeof(T) * (size_t)c.N, hipMemcpyHostToDevice, ctx->stream));
    k_code<T><<<nblocks(c.N), BLOCK, 0, ctx->stream>>>((const T*)ctx->d_f, d, c, ctx->d_code_f, nullptr);
    CK(hipGetLastError());
    int64_t* outs[4] = {pv, pe, pf, pt};
    i64 sizes[4] = {c.N, c.E, c.F, c.T};
    for (int k = 0; k < 4; k++) {
        int64_t* dbuf = nullptr;
        CK(hipMalloc(&dbuf, sizeof(int64_t) * (size_t)(sizes[k] > 0 ? sizes[k] : 1)));
        if (sizes[k] > 0) {
            k_pairing<T><<<nblocks(sizes[k]), BLOCK, 0, ctx->stream>>>((const T*)ctx->d_f, ctx->d_code_f, d, c, k,
                                                                        (ull)sizes[k], dbuf);
            hipError_t e = hipGetLastError();
            if (e == hipSuccess)
                e = hipMemcpyAsync(outs[k], dbuf, sizeof(int64_t) * (size_t)sizes[k], hipMemcpyDeviceToHost, ctx->stream);
            if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
            (void)hipFree(dbuf);
            if (e != hipSuccess) {
                ctx->err = hipGetErrorString(e);
                return DMTZ_E_HIP;
            }
        } else {
            (void)hipFree(dbuf);
        }
    }
    return 0;
}

template <class T>
int targets_host(dmtz_ctx* ctx, const T* f, const T* g, i64 nx, i64 ny, i64 nz, double eps, double range,
                 const dmtz_flags* flags, uint8_t* targets, int64_t* counts4) {
    int rc = check_dims(ctx, nx, ny, nz);
    if (rc) return rc;
    Dims d{nx, ny, nz};
    Counts c = make_counts(d);
    rc = ensure(ctx, c.N, sizeof(T), true);
    if (rc) return rc;
    CK(hipMemcpyAsync(ctx->d_f, f, sizeof(T) * (size_t)c.N, hipMemcpyHostToDevice, ctx->stream));
    CK(hipMemcpyAsync(ctx->d_g, g, sizeof(T) * (size_t)c.N, hipMemcpyHostToDevice, ctx->stream));
    CK(hipMemsetAsync(ctx->d_counter, 1, (size_t)c.N, ctx->stream));
    CK(hipMemsetAsync(ctx->d_ctr, 0, sizeof(ull) * K_NCTR, ctx->stream));
    Driver<T> drv{ctx, (const T*)ctx->d_f, (T*)ctx->d_g, d, c, eps, range, *flags};
    rc = drv.f_codes();
    if (rc) return rc;
    rc = drv.g_codes();
    if (rc) return rc;
    rc = drv.detect();
    if (rc) return rc;
    ull cc[4] = {ctx->h_ctr[0], ctx->h_ctr[1], ctx->h_ctr[2], ctx->h_ctr[3]};
    for (int k = 0; k < 4; k++) counts4[k] = (int64_t)cc[k];
    rc = drv.begin_edits();
    if (rc) return rc;
    ull nrec = 0;
    rc = drv.repair_records(cc, nrec);
    if (rc) return rc;
    rc = sync_ctr(ctx);
    if (rc) return rc;
    ull nt = ctx->h_ctr[K_TGT];
    std::vector<uint32_t> tl(nt);
    if (nt) CK(hipMemcpyAsync(tl.data(), ctx->d_targets, sizeof(uint32_t) * nt, hipMemcpyDeviceToHost, ctx->stream));
    CK(hipStreamSynchronize(ctx->stream));
    std::memset(targets, 0, (size_t)c.N);
    for (ull k = 0; k < nt; k++) targets[tl[k]] = 1;
    return 0;
}

template <class T>
int paths_host(dmtz_ctx* ctx, const T* f, const T* g, i64 nx, i64 ny, i64 nz, const dmtz_flags* flags,
               int64_t* counts2, int64_t* ds_id, int64_t* ds_rec, int64_t* as_id, int64_t* as_rec, int64_t cap,
               uint8_t* targets) {
    int rc = check_dims(ctx, nx, ny, nz);
    if (rc) return rc;
    if (!flags || !counts2) {
        ctx->err = "null argument";
        return DMTZ_E_ARG;
    }
    Dims d{nx, ny, nz};
    Counts c = make_counts(d);
    rc = ensure(ctx, c.N, sizeof(T), true);
    if (rc) return rc;
    CK(hipMemcpyAsync(ctx->d_f, f, sizeof(T) * (size_t)c.N, hipMemcpyHostToDevice, ctx->stream));
    CK(hipMemcpyAsync(ctx->d_g, g, sizeof(T) * (size_t)c.N, hipMemcpyHostToDevice, ctx->stream));
    CK(hipMemsetAsync(ctx->d_counter, 1, (size_t)c.N, ctx->stream));
    CK(hipMemsetAsync(ctx->d_ctr, 0, sizeof(ull) * K_NCTR, ctx->stream));
    dmtz_flags fl = *flags;
    fl.preserve_vpath = 1;
    fl.preserve_connectors = 0;
    Driver<T> drv{ctx, (const T*)ctx->d_f, (T*)ctx->d_g, d, c, 0.0, 0.0, fl};
    rc = drv.f_codes();
    if (rc) return rc;
    rc = drv.g_codes();
    if (rc) return rc;
    uint64_t* gfaces = nullptr;
    if (fl.preserve_geometry != 1) {
        // get_false_as_vpath_connectivity starts from g's critical faces (main_3d.cu:4766)
        rc = drv.fcrit_of(ctx->d_code_g, 4);
        if (rc) return rc;
        drv.n_gfaces = ctx->h_ctr[K_FC0 + 2];
        CK(hipMalloc(&gfaces, sizeof(uint64_t) * (drv.n_gfaces + 1)));
        hipError_t e = hipMemcpyAsync(gfaces, fcrit_lists(ctx).l[2], sizeof(uint64_t) * drv.n_gfaces,
                                      hipMemcpyDeviceToDevice, ctx->stream);
        if (e != hipSuccess) {
            (void)hipFree(gfaces);
            ctx->err = hipGetErrorString(e);
            return DMTZ_E_HIP;
        }
        drv.as_gfaces = gfaces;
    }
    rc = drv.prepare_paths();
    if (rc == 0) rc = drv.detect_paths();
    if (gfaces) {
        (void)hipStreamSynchronize(ctx->stream);
        (void)hipFree(gfaces);
    }
    if (rc) return rc;
    ull nds = ctx->h_ctr[K_DS], nas = ctx->h_ctr[K_AS];
    counts2[0] = (int64_t)nds;
    counts2[1] = (int64_t)nas;
    PathList pd = drv.path_list(0), pa = drv.path_list(1);
    ull kd = nds < (ull)cap ? nds : (ull)cap, ka = nas < (ull)cap ? nas : (ull)cap;
    if (kd && ds_id) CK(hipMemcpyAsync(ds_id, pd.id, sizeof(int64_t) * kd, hipMemcpyDeviceToHost, ctx->stream));
    if (kd && ds_rec) CK(hipMemcpyAsync(ds_rec, pd.rec, sizeof(int64_t) * kd, hipMemcpyDeviceToHost, ctx->stream));
    if (ka && as_id) CK(hipMemcpyAsync(as_id, pa.id, sizeof(int64_t) * ka, hipMemcpyDeviceToHost, ctx->stream));
    if (ka && as_rec) CK(hipMemcpyAsync(as_rec, pa.rec, sizeof(int64_t) * ka, hipMemcpyDeviceToHost, ctx->stream));
    CK(hipStreamSynchronize(ctx->stream));
    if (targets) {
        rc = drv.begin_edits();
        if (rc) return rc;
        if (nds + nas) {
            k_fix_paths<T><<<nblocks(nds + nas), BLOCK, 0, ctx->stream>>>(
                pa.rec, nas, pd.rec, nds, nullptr, 0, (const T*)ctx->d_f, ctx->d_code_f, (const T*)ctx->d_g,
                ctx->d_code_g, d, c, ctx->d_counter, 0.0, 0.0, fl.edit_type, drv.sink());
            CK(hipGetLastError());
        }
        rc = sync_ctr(ctx);
        if (rc) return rc;
        ull nt = ctx->h_ctr[K_TGT];
        std::vector<uint32_t> tl(nt);
        if (nt) CK(hipMemcpyAsync(tl.data(), ctx->d_targets, sizeof(uint32_t) * nt, hipMemcpyDeviceToHost, ctx->stream));
        CK(hipStreamSynchronize(ctx->stream));
        std::memset(targets, 0, (size_t)c.N);
        for (ull k = 0; k < nt; k++) targets[tl[k]] = 1;
    }
    CK(hipMemcpyAsync(ctx->h_ctr, ctx->d_ctr, sizeof(ull) * K_NCTR, hipMemcpyDeviceToHost, ctx->stream));
    CK(hipStreamSynchronize(ctx->stream));
    if (ctx->h_ctr[K_WALK] > 0) {
        ctx->err = "a separatrix walk exceeded N steps (cyclic gradient path)";
        return DMTZ_E_WALK;
    }
    return 0;
}

// The 2-saddle wall list on (f, g) as they stand, and optionally the edit set of fix_ds_wall over
// it (eps and range enter its threshold, main_3d.cu:5331).
template <class T>
int walls_host(dmtz_ctx* ctx, const T* f, const T* g, i64 nx, i64 ny, i64 nz, double eps, double range,
               const dmtz_flags* flags, int64_t* count, int64_t* face, int64_t* edge, int64_t cap, uint8_t* targets) {
    int rc = check_dims(ctx, nx, ny, nz);
    if (rc) return rc;
    if (!count || cap < 0) {
        ctx->err = "null argument";
        return DMTZ_E_ARG;
    }
    Dims d{nx, ny, nz};
    Counts c = make_counts(d);
    rc = ensure(ctx, c.N, sizeof(T), true);
    if (rc) return rc;
    CK(hipMemcpyAsync(ctx->d_f, f, sizeof(T) * (size_t)c.N, hipMemcpyHostToDevice, ctx->stream));
    CK(hipMemcpyAsync(ctx->d_g, g, sizeof(T) * (size_t)c.N, hipMemcpyHostToDevice, ctx->stream));
    CK(hipMemsetAsync(ctx->d_counter, 1, (size_t)c.N, ctx->stream));
    CK(hipMemsetAsync(ctx->d_ctr, 0, sizeof(ull) * K_NCTR, ctx->stream));
    dmtz_flags fl{};
    if (flags) fl = *flags;
    fl.preserve_vpath = 1;
    fl.preserve_connectors = 1;
    Driver<T> drv{ctx, (const T*)ctx->d_f, (T*)ctx->d_g, d, c, eps, range, fl};
    rc = drv.f_codes();
    if (rc) return rc;
    rc = drv.g_codes();
    if (rc) return rc;
    rc = drv.prepare_paths();
    if (rc) return rc;
    rc = drv.walls();
    if (rc) return rc;
    rc = sync_ctr(ctx);
    if (rc) return rc;
    rc = overflow_check(ctx);
    if (rc) return rc;
    const ull nw = ctx->h_ctr[K_DSW];
    *count = (int64_t)nw;
    PathList pw = drv.path_list(2);
    const ull k = nw < (ull)cap ? nw : (ull)cap;
    if (k && face) CK(hipMemcpyAsync(face, pw.id, sizeof(int64_t) * k, hipMemcpyDeviceToHost, ctx->stream));
    if (k && edge) CK(hipMemcpyAsync(edge, pw.rec, sizeof(int64_t) * k, hipMemcpyDeviceToHost, ctx->stream));
    CK(hipStreamSynchronize(ctx->stream));
    if (targets) {
        rc = drv.begin_edits();
        if (rc) return rc;
        if (nw) {
            k_fix_paths<T><<<nblocks(nw), BLOCK, 0, ctx->stream>>>(
                nullptr, 0, nullptr, 0, pw.rec, nw, (const T*)ctx->d_f, ctx->d_code_f, (const T*)ctx->d_g,
                ctx->d_code_g, d, c, ctx->d_counter, eps, range, fl.edit_type, drv.sink());
            CK(hipGetLastError());
        }
        rc = sync_ctr(ctx);
        if (rc) return rc;
        const ull nt = ctx->h_ctr[K_TGT];
        std::vector<uint32_t> tl(nt);
        if (nt) CK(hipMemcpyAsync(tl.data(), ctx->d_targets, sizeof(uint32_t) * nt, hipMemcpyDeviceToHost, ctx->stream));
        CK(hipStreamSynchronize(ctx->stream));
        std::memset(targets, 0, (size_t)c.N);
        for (ull q = 0; q < nt; q++) targets[tl[q]] = 1;
    }
    return 0;
}

template <class T>
int eps_from_rel(const T* f, int64_t n, double rel, double* eps_out, double* range_out) {
    if (!f || n <= 0) return DMTZ_E_ARG;
    double lo = (double)f[0], hi = (double)f[0];
    for (int64_t i = 1; i < n; i++) {
        double x = (double)f[i];
        lo = x < lo ? x : lo;
        hi = x > hi ? x : hi;
    }
    double range = hi - lo;
    if (range_out) *range_out = range;
    if (eps_out) *eps_out = range * rel;  // bound1 = (maxValue - minValue) * range1, :6307
    return 0;
}

}  // namespace
