// The edit stream of a correction (include/dmtz.h dmtz_edit_stream_* / dmtz_edit_apply_*): the
// reference's post-hoc encodings of what the correction changed, built and decoded on the device.
//
//   edit_type 1 (cost, main_3d.cu:1494-1646): a 3-bit index per vertex, packed least significant
//     bit first into bytes (encodeTo3BitBitmap, main_3d.cu:1446-1476): 0 = unchanged, k = 1..5 =
//     changed by the k - 1 steps -2^-s eps (s = 1 .. k-1) of the counter-driven edit
//     (main_3d.cu:241, 1131-1137), 6 = its value is carried explicitly; plus one double per
//     index 6, in vertex order: the change g' - g (main_3d.cu:1502-1508, where it reads
//     (f - eps) - g, the same double for the reference's snap).
//   edit_type 0 (original_cost, main_3d.cu:1649-1757): the changed vertices' indices as int32
//     differences (the first index, then the gap to the previous one), plus g' - g per changed
//     vertex as a double.
//
// A changed vertex whose counter does not replay to its value (a counter reset by filtration,
// main_3d.cu:6782, a learned warm edit, or the finisher's restore, counter 7) is written with
// index 6 and its value, so every stream decodes back to the corrected field.
//
// Kernels: a mark pass (index per vertex and a flag for the vertices that go to the value
// stream), an exclusive scan of the flags (vertex order is kept), a write pass. The bitmap byte j
// gathers the indices of vertices floor(8j/3) .. floor((8j+7)/3).
#include "driver.h"

namespace {

constexpr int CODEC_BLOCK = 256;

inline unsigned codec_blocks(i64 n) { return (unsigned)((n + CODEC_BLOCK - 1) / CODEC_BLOCK); }

// The value the counter-driven edit reaches from g0 after k - 1 accepted steps (the candidate
// arithmetic of EditArith<T>::step without its bound test, which every accepted step passed).
template <class T>
__device__ __forceinline__ T replay_steps(T g0, int k, double eps) {
    T g = g0;
    for (int s = 1; s < k; s++) g = (T)((double)g + halving_delta(eps, s));
    return g;
}

// Index and value-stream flag of every vertex; *edited counts the changed vertices.
template <class T>
__global__ __launch_bounds__(CODEC_BLOCK) void k_es_mark(const T* g0, const T* g1, const uint8_t* counter, i64 n,
                                                         double eps, int edit_type, uint8_t* idx, uint32_t* flag,
                                                         ull* edited) {
    const i64 i = (i64)blockIdx.x * CODEC_BLOCK + threadIdx.x;
    bool changed = false;
    if (i < n) {
        const T a = g0[i], b = g1[i];
        uint8_t k = 0;
        uint32_t fl = 0;
        changed = !(a == b);
        if (changed) {
            if (edit_type == 1) {
                const int c = counter[i];
                k = (c >= 1 && c <= 5 && replay_steps<T>(a, c, eps) == b) ? (uint8_t)c : (uint8_t)6;
                fl = k == 6;
            } else {
                k = 1;
                fl = 1;
            }
        }
        idx[i] = k;
        flag[i] = fl;
    }
    wave_add(edited, changed);
}

__global__ __launch_bounds__(CODEC_BLOCK) void k_es_pack3(const uint8_t* idx, i64 n, i64 nbytes, uint8_t* out) {
    const i64 j = (i64)blockIdx.x * CODEC_BLOCK + threadIdx.x;
    if (j >= nbytes) return;
    const i64 lo = 8 * j, hi = 8 * j + 8;  // stream bits of this byte
    uint32_t byte = 0;
    for (i64 v = lo / 3; v * 3 < hi && v < n; v++) {
        const i64 sh = 3 * v - lo;  // position of the vertex's bit 0 relative to the byte
        const uint32_t x = idx[v] & 7u;
        byte |= sh >= 0 ? (x << sh) : (x >> (-sh));
    }
    out[j] = (uint8_t)(byte & 0xFFu);
}

template <class T>
__global__ __launch_bounds__(CODEC_BLOCK) void k_es_values(const T* g0, const T* g1, const uint32_t* flag,
                                                           const uint32_t* pos, i64 n, int edit_type, double* edits,
                                                           int32_t* ids) {
    const i64 i = (i64)blockIdx.x * CODEC_BLOCK + threadIdx.x;
    if (i >= n || !flag[i]) return;
    const uint32_t p = pos[i];
    edits[p] = (double)g1[i] - (double)g0[i];
    if (edit_type == 0) ids[p] = (int32_t)i;
}

__global__ __launch_bounds__(CODEC_BLOCK) void k_es_diffs(const int32_t* ids, i64 m, int32_t* out) {
    const i64 p = (i64)blockIdx.x * CODEC_BLOCK + threadIdx.x;
    if (p >= m) return;
    out[p] = p == 0 ? ids[0] : ids[p] - ids[p - 1];
}

// Decoding, edit_type 1: index of vertex i from the bitmap.
__device__ __forceinline__ int es_index3(const uint8_t* bm, i64 i) {
    const i64 bit = 3 * i;
    const i64 j = bit >> 3;
    const int sh = (int)(bit & 7);
    uint32_t w = bm[j];
    if (sh > 5) w |= (uint32_t)bm[j + 1] << 8;
    return (int)((w >> sh) & 7u);
}

__global__ __launch_bounds__(CODEC_BLOCK) void k_ed_flags3(const uint8_t* bm, i64 n, uint32_t* flag, uint32_t* bad) {
    const i64 i = (i64)blockIdx.x * CODEC_BLOCK + threadIdx.x;
    if (i >= n) return;
    const int k = es_index3(bm, i);
    flag[i] = k == 6;
    if (k == 7) atomicOr(bad, 1u);
}

template <class T>
__global__ __launch_bounds__(CODEC_BLOCK) void k_ed_apply3(const uint8_t* bm, const uint32_t* pos, i64 n, double eps,
                                                           const double* edits, T* g) {
    const i64 i = (i64)blockIdx.x * CODEC_BLOCK + threadIdx.x;
    if (i >= n) return;
    const int k = es_index3(bm, i);
    if (k == 6) g[i] = (T)((double)g[i] + edits[pos[i]]);
    else if (k >= 2 && k <= 5) g[i] = replay_steps<T>(g[i], k, eps);
}

// Decoding, edit_type 0: absolute indices are the inclusive sums of the differences.
template <class T>
__global__ __launch_bounds__(CODEC_BLOCK) void k_ed_apply0(const int64_t* ids, i64 m, i64 n, const double* edits, T* g,
                                                           uint32_t* bad) {
    const i64 p = (i64)blockIdx.x * CODEC_BLOCK + threadIdx.x;
    if (p >= m) return;
    const i64 i = ids[p];
    if (i < 0 || i >= n) {
        atomicOr(bad, 1u);
        return;
    }
    g[i] = (T)((double)g[i] + edits[p]);
}

__global__ __launch_bounds__(CODEC_BLOCK) void k_widen(const int32_t* a, i64 m, int64_t* out) {
    const i64 p = (i64)blockIdx.x * CODEC_BLOCK + threadIdx.x;
    if (p < m) out[p] = a[p];
}

// Run-length code of a bitmap (runLengthEncodeBitmap main_3d.cu:1399-1426): the bits of every
// byte from the most significant down. Bit k of the stream is bit 7 - (k mod 8) of byte k / 8, so a
// byte read as an 8-bit word already lists its bits in stream order, and the word of "previous
// bits" is the byte shifted right by one with the previous byte's last bit (its bit 0) on top. A
// run starts where the two differ, and at bit 0 of the stream.
__device__ __forceinline__ uint32_t rle_starts(const uint8_t* bm, i64 j) {
    const uint32_t x = bm[j];
    const uint32_t prev = j ? (uint32_t)(bm[j - 1] & 1u) : (~x >> 7) & 1u;
    return (x ^ ((x >> 1) | (prev << 7))) & 0xFFu;
}

__global__ __launch_bounds__(CODEC_BLOCK) void k_rle_count(const uint8_t* bm, i64 nbytes, uint32_t* cnt) {
    const i64 j = (i64)blockIdx.x * CODEC_BLOCK + threadIdx.x;
    if (j < nbytes) cnt[j] = (uint32_t)__popc(rle_starts(bm, j));
}

// Run r's bit into pairs[2r] and its first stream position into start[r].
__global__ __launch_bounds__(CODEC_BLOCK) void k_rle_starts(const uint8_t* bm, i64 nbytes, const uint32_t* pos,
                                                            int32_t* pairs, uint32_t* start) {
    const i64 j = (i64)blockIdx.x * CODEC_BLOCK + threadIdx.x;
    if (j >= nbytes) return;
    uint32_t s = rle_starts(bm, j);
    const uint32_t x = bm[j];
    uint32_t r = pos[j];
    while (s) {
        const int b = 31 - __clz(s);  // highest set bit first: stream order
        s &= ~(1u << b);
        pairs[2 * (i64)r] = (int32_t)((x >> b) & 1u);
        start[r] = (uint32_t)(8 * j + (7 - b));
        r++;
    }
}

__global__ __launch_bounds__(CODEC_BLOCK) void k_rle_len(const uint32_t* start, i64 runs, i64 nbits, int32_t* pairs) {
    const i64 r = (i64)blockIdx.x * CODEC_BLOCK + threadIdx.x;
    if (r >= runs) return;
    const i64 end = r + 1 < runs ? (i64)start[r + 1] : nbits;
    pairs[2 * r + 1] = (int32_t)(end - (i64)start[r]);
}

// Scratch of the codec: per-vertex index bytes, flags and their scan, scan storage.
struct CodecScratch {
    dmtz_ctx* ctx;
    uint8_t* idx = nullptr;
    uint32_t* flag = nullptr;
    uint32_t* pos = nullptr;
    void* tmp = nullptr;
    int32_t* ids = nullptr;
    int64_t* ids64 = nullptr;
    uint32_t* bad = nullptr;  // error flag (decoding) or changed-vertex counter (encoding, 8 bytes)
    ~CodecScratch() {
        if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
        (void)hipFree(idx);
        (void)hipFree(flag);
        (void)hipFree(pos);
        (void)hipFree(tmp);
        (void)hipFree(ids);
        (void)hipFree(ids64);
        (void)hipFree(bad);
    }
};

// Exclusive sum of n uint32 flags into pos; returns the total through *total.
int scan_flags(dmtz_ctx* ctx, CodecScratch& S, const uint32_t* flag, uint32_t* pos, i64 n, ull* total) {
    size_t tb = 0;
    CK(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, flag, pos, (int)n, ctx->stream));
    CK(hipMalloc(&S.tmp, tb + 16));
    CK(hipcub::DeviceScan::ExclusiveSum(S.tmp, tb, flag, pos, (int)n, ctx->stream));
    uint32_t last_pos = 0, last_flag = 0;
    CK(hipMemcpyAsync(&last_pos, pos + (n - 1), sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream));
    CK(hipMemcpyAsync(&last_flag, flag + (n - 1), sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream));
    CK(hipStreamSynchronize(ctx->stream));
    *total = (ull)last_pos + last_flag;
    return 0;
}

template <class T>
int edit_stream(dmtz_ctx* ctx, const T* d_g0, const T* d_g1, const uint8_t* d_counter, int64_t n, double eps,
                int32_t edit_type, uint8_t* d_index, int64_t index_cap, double* d_edits, int64_t edits_cap,
                dmtz_edit_stream* out) {
    if (!ctx) return DMTZ_E_ARG;
    if (!d_g0 || !d_g1 || !out || n <= 0 || n >= ((int64_t)1 << 31) || (edit_type != 0 && edit_type != 1) ||
        (edit_type == 1 && !d_counter)) {
        ctx->err = "edit stream: null or invalid argument";
        return DMTZ_E_ARG;
    }
    CK(hipSetDevice(ctx->device));
    CodecScratch S{ctx};
    CK(hipMalloc(&S.idx, (size_t)n));
    CK(hipMalloc(&S.flag, sizeof(uint32_t) * (size_t)n));
    CK(hipMalloc(&S.pos, sizeof(uint32_t) * (size_t)n));
    CK(hipMalloc(&S.bad, sizeof(ull)));
    CK(hipMemsetAsync(S.bad, 0, sizeof(ull), ctx->stream));
    k_es_mark<T><<<codec_blocks(n), CODEC_BLOCK, 0, ctx->stream>>>(d_g0, d_g1, d_counter, n, eps, edit_type, S.idx,
                                                                     S.flag, (ull*)S.bad);
    CK(hipGetLastError());
    ull nvals = 0;
    int rc = scan_flags(ctx, S, S.flag, S.pos, n, &nvals);
    if (rc) return rc;
    ull edited = 0;
    CK(hipMemcpyAsync(&edited, S.bad, sizeof(ull), hipMemcpyDeviceToHost, ctx->stream));
    CK(hipStreamSynchronize(ctx->stream));
    const int64_t index_bytes = edit_type == 1 ? (3 * n + 7) / 8 : (int64_t)(4 * edited);
    out->edited = (int64_t)edited;
    out->index_bytes = index_bytes;
    out->n_edits = (int64_t)nvals;
    if (!d_index || !d_edits || index_cap < index_bytes || edits_cap < (int64_t)nvals) {
        ctx->err = "edit stream: output buffers smaller than the sizes returned";
        return DMTZ_E_CAPACITY;
    }
    if (edit_type == 0) CK(hipMalloc(&S.ids, sizeof(int32_t) * (size_t)(nvals + 1)));
    k_es_values<T><<<codec_blocks(n), CODEC_BLOCK, 0, ctx->stream>>>(d_g0, d_g1, S.flag, S.pos, n, edit_type, d_edits,
                                                                       S.ids);
    CK(hipGetLastError());
    if (edit_type == 1) {
        k_es_pack3<<<codec_blocks(index_bytes), CODEC_BLOCK, 0, ctx->stream>>>(S.idx, n, index_bytes, d_index);
        CK(hipGetLastError());
    } else if (nvals) {
        k_es_diffs<<<codec_blocks((i64)nvals), CODEC_BLOCK, 0, ctx->stream>>>(S.ids, (i64)nvals, (int32_t*)d_index);
        CK(hipGetLastError());
    }
    CK(hipStreamSynchronize(ctx->stream));
    return 0;
}

template <class T>
int edit_apply(dmtz_ctx* ctx, T* d_g, int64_t n, double eps, int32_t edit_type, const uint8_t* d_index,
               int64_t index_bytes, const double* d_edits, int64_t n_edits) {
    if (!ctx) return DMTZ_E_ARG;
    if (!d_g || n <= 0 || n >= ((int64_t)1 << 31) || (edit_type != 0 && edit_type != 1) || n_edits < 0 ||
        (n_edits > 0 && (!d_edits || !d_index))) {
        ctx->err = "edit apply: null or invalid argument";
        return DMTZ_E_ARG;
    }
    if ((edit_type == 1 && (index_bytes != (3 * n + 7) / 8 || !d_index)) ||
        (edit_type == 0 && (index_bytes != 4 * n_edits || n_edits > n))) {
        ctx->err = "edit apply: index stream size does not match the field and the value stream";
        return DMTZ_E_ARG;
    }
    CK(hipSetDevice(ctx->device));
    CodecScratch S{ctx};
    CK(hipMalloc(&S.bad, sizeof(uint32_t)));
    CK(hipMemsetAsync(S.bad, 0, sizeof(uint32_t), ctx->stream));
    if (edit_type == 1) {
        CK(hipMalloc(&S.flag, sizeof(uint32_t) * (size_t)n));
        CK(hipMalloc(&S.pos, sizeof(uint32_t) * (size_t)n));
        k_ed_flags3<<<codec_blocks(n), CODEC_BLOCK, 0, ctx->stream>>>(d_index, n, S.flag, S.bad);
        CK(hipGetLastError());
        ull total = 0;
        int rc = scan_flags(ctx, S, S.flag, S.pos, n, &total);
        if (rc) return rc;
        if ((int64_t)total != n_edits) {
            ctx->err = "edit apply: the bitmap's index-6 count differs from the value stream's length";
            return DMTZ_E_ARG;
        }
        k_ed_apply3<T><<<codec_blocks(n), CODEC_BLOCK, 0, ctx->stream>>>(d_index, S.pos, n, eps, d_edits, d_g);
        CK(hipGetLastError());
    } else if (n_edits) {
        CK(hipMalloc(&S.ids64, sizeof(int64_t) * (size_t)n_edits));
        k_widen<<<codec_blocks(n_edits), CODEC_BLOCK, 0, ctx->stream>>>((const int32_t*)d_index, n_edits, S.ids64);
        CK(hipGetLastError());
        size_t tb = 0;
        CK(hipcub::DeviceScan::InclusiveSum(nullptr, tb, S.ids64, S.ids64, (int)n_edits, ctx->stream));
        CK(hipMalloc(&S.tmp, tb + 16));
        CK(hipcub::DeviceScan::InclusiveSum(S.tmp, tb, S.ids64, S.ids64, (int)n_edits, ctx->stream));
        k_ed_apply0<T><<<codec_blocks(n_edits), CODEC_BLOCK, 0, ctx->stream>>>(S.ids64, n_edits, n, d_edits, d_g,
                                                                                S.bad);
        CK(hipGetLastError());
    }
    uint32_t bad = 0;
    CK(hipMemcpyAsync(&bad, S.bad, sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream));
    CK(hipStreamSynchronize(ctx->stream));
    if (bad) {
        ctx->err = "edit apply: index stream holds an index outside the field or the value 7";
        return DMTZ_E_ARG;
    }
    return 0;
}

int bitmap_rle(dmtz_ctx* ctx, const uint8_t* d_bitmap, int64_t nbytes, int32_t* d_pairs, int64_t cap_runs,
               int64_t* n_runs) {
    if (!ctx) return DMTZ_E_ARG;
    if (!n_runs || nbytes < 0 || (nbytes > 0 && !d_bitmap) || nbytes > (int64_t)(INT32_MAX / 8)) {
        ctx->err = "bitmap rle: null or invalid argument (at most 2^28 - 1 bytes)";
        return DMTZ_E_ARG;
    }
    *n_runs = 0;
    if (nbytes == 0) return 0;
    CK(hipSetDevice(ctx->device));
    CodecScratch S{ctx};
    CK(hipMalloc(&S.flag, sizeof(uint32_t) * (size_t)nbytes));
    CK(hipMalloc(&S.pos, sizeof(uint32_t) * (size_t)nbytes));
    k_rle_count<<<codec_blocks(nbytes), CODEC_BLOCK, 0, ctx->stream>>>(d_bitmap, nbytes, S.flag);
    CK(hipGetLastError());
    ull runs = 0;
    int rc = scan_flags(ctx, S, S.flag, S.pos, nbytes, &runs);
    if (rc) return rc;
    *n_runs = (int64_t)runs;
    if (!d_pairs || cap_runs < (int64_t)runs) {
        ctx->err = "bitmap rle: output buffer smaller than the run count returned";
        return DMTZ_E_CAPACITY;
    }
    CK(hipMalloc(&S.ids, sizeof(uint32_t) * (size_t)runs));
    k_rle_starts<<<codec_blocks(nbytes), CODEC_BLOCK, 0, ctx->stream>>>(d_bitmap, nbytes, S.pos, d_pairs,
                                                                        (uint32_t*)S.ids);
    CK(hipGetLastError());
    k_rle_len<<<codec_blocks((i64)runs), CODEC_BLOCK, 0, ctx->stream>>>((const uint32_t*)S.ids, (i64)runs, 8 * nbytes,
                                                                       d_pairs);
    CK(hipGetLastError());
    CK(hipStreamSynchronize(ctx->stream));
    return 0;
}

}  // namespace

extern "C" {

int dmtz_bitmap_rle(dmtz_ctx* ctx, const uint8_t* d_bitmap, int64_t nbytes, int32_t* d_pairs, int64_t cap_runs,
                    int64_t* n_runs) {
    return bitmap_rle(ctx, d_bitmap, nbytes, d_pairs, cap_runs, n_runs);
}

int dmtz_edit_stream_f32(dmtz_ctx* ctx, const float* d_g0, const float* d_g1, const uint8_t* d_counter, int64_t n,
                         double eps, int32_t edit_type, uint8_t* d_index, int64_t index_cap, double* d_edits,
                         int64_t edits_cap, dmtz_edit_stream* out) {
    return edit_stream<float>(ctx, d_g0, d_g1, d_counter, n, eps, edit_type, d_index, index_cap, d_edits, edits_cap,
                              out);
}
int dmtz_edit_stream_f64(dmtz_ctx* ctx, const double* d_g0, const double* d_g1, const uint8_t* d_counter, int64_t n,
                         double eps, int32_t edit_type, uint8_t* d_index, int64_t index_cap, double* d_edits,
                         int64_t edits_cap, dmtz_edit_stream* out) {
    return edit_stream<double>(ctx, d_g0, d_g1, d_counter, n, eps, edit_type, d_index, index_cap, d_edits, edits_cap,
                               out);
}
int dmtz_edit_apply_f32(dmtz_ctx* ctx, float* d_g, int64_t n, double eps, int32_t edit_type, const uint8_t* d_index,
                        int64_t index_bytes, const double* d_edits, int64_t n_edits) {
    return edit_apply<float>(ctx, d_g, n, eps, edit_type, d_index, index_bytes, d_edits, n_edits);
}
int dmtz_edit_apply_f64(dmtz_ctx* ctx, double* d_g, int64_t n, double eps, int32_t edit_type, const uint8_t* d_index,
                        int64_t index_bytes, const double* d_edits, int64_t n_edits) {
    return edit_apply<double>(ctx, d_g, n, eps, edit_type, d_index, index_bytes, d_edits, n_edits);
}

}  // extern "C"
