// The learned warm start (include/dmtz_learned.h): a four-layer 3x3x3 convolutional network over
// 16-channel bf16 activations, run on the matrix cores, trained by the library's own backward
// kernels and Adam. No counterpart in the reference.
//
// Convolutions are implicit GEMMs with K = 16 channels per tap: a wave computes a 16-voxel run
// along x for all 16 output channels with one v_mfma_f32_16x16x16_bf16 per tap (27 per run).
// Operand layout of that instruction (per lane l, i = l % 16, q = l / 16, r = 0..3):
//   A[row i][k = 4q + r], B[k = 4q + r][col i], D[row 4q + r][col i].
// Activations are stored channel-last, 16 bf16 (32 bytes) per vertex, so the A operand of a
// forward tap is one 8-byte load per lane. Weights of a layer (27 x 16 x 16 bf16, 13.5 KiB) are
// staged in LDS per block.
//
// Kernels:
//   k_features   input channels from g (lower-star critical cells and neighbourhood statistics)
//   k_conv<0|1|2> forward hidden layer (bias, ReLU), forward output layer (bias, fp32 out),
//                backward data (transposed taps and weights, times ReLU' of the layer input)
//   k_wgrad      weight and bias gradients (MFMA over voxel runs, LDS then global reduction)
//   k_loss       loss terms and the output gradient
//   k_adam       Adam on the fp32 master parameters, bf16 copies of the weights
//   k_warm       the warm-start edit of g
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <random>
#include <vector>

#include "../../include/dmtz.h"
#include "../../include/dmtz_learned.h"
#include "edit.h"
#include "mismatch.h"

namespace dmtz_internal {
hipStream_t stream(dmtz_ctx* ctx);
int device(dmtz_ctx* ctx);
void set_error(dmtz_ctx* ctx, const char* msg);
}  // namespace dmtz_internal

using namespace dmtz;

namespace {

typedef short short4_t __attribute__((ext_vector_type(4)));
typedef float float4_t __attribute__((ext_vector_type(4)));

constexpr int CH = DMTZ_NET_CHANNELS;    // channels per vertex (padded)
constexpr int TAPS = 27;
constexpr int WSZ = TAPS * CH * CH;      // weights of one layer
constexpr int LP = DMTZ_NET_LAYER_PARAMS;  // WSZ + CH
constexpr int NL = 4;
constexpr int CIN0 = 8;                  // used input channels
constexpr int COUT_LAST = 2;             // used output channels
constexpr int BLOCK = 256;
constexpr int TXC = 18;                  // a 16-voxel run and its halo along x

static_assert(LP == WSZ + CH, "layer parameter layout");

// bf16 <-> fp32 (round to nearest even; inputs are finite)
__host__ __device__ inline uint16_t f2bf(float x) {
    uint32_t u;
    memcpy(&u, &x, 4);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__host__ __device__ inline float bf2f(uint16_t h) {
    const uint32_t u = (uint32_t)h << 16;
    float x;
    memcpy(&x, &u, 4);
    return x;
}

__device__ inline void tap_offset(int tap, int& dx, int& dy, int& dz) {
    dx = tap % 3 - 1;
    dy = (tap / 3) % 3 - 1;
    dz = tap / 9 - 1;
}

// Run w (16 voxels along x) of a grid with nxs runs per row.
__device__ inline void run_origin(long w, int nxs, int ny, int& x0, int& y, int& z) {
    x0 = (int)(w % nxs) * 16;
    const long rest = w / nxs;
    y = (int)(rest % ny);
    z = (int)(rest / ny);
}

// Input channels of vertex v from g.
__global__ __launch_bounds__(BLOCK) void k_features(const float* g, Dims d, i64 n, float inv_eps, float eps_ch,
                                                    uint16_t* feat) {
    const i64 v = (i64)blockIdx.x * BLOCK + threadIdx.x;
    if (v >= n) return;
    LowerStar ls;
    star_of(g, d, v, ls);
    StarCode sc;
    encode_star(ls, sc);
    uint64_t clo, ilo;
    uint32_t chi, ihi;
    code_masks(sc, clo, chi, ilo, ihi);
    const uint64_t cl = clo & ilo;
    const uint32_t ch = chi & ihi;
    i64 x, y, z;
    vxyz(d, v, x, y, z);
    const float gv = g[v];
    float sum = 0.f, mn = gv, mx = gv;
    int cnt = 0;
    for (int k = 0; k < 14; k++) {
        const i64 xx = x + DIR[k][0], yy = y + DIR[k][1], zz = z + DIR[k][2];
        if (!inb(d, xx, yy, zz)) continue;
        const float w = g[vidx(d, xx, yy, zz)];
        sum += w;
        mn = w < mn ? w : mn;
        mx = w > mx ? w : mx;
        cnt++;
    }
    auto clampf = [](float a) { return a < -8.f ? -8.f : (a > 8.f ? 8.f : a); };
    float c[CH];
    for (int k = 0; k < CH; k++) c[k] = 0.f;
    c[0] = clampf((gv - (cnt ? sum / (float)cnt : gv)) * inv_eps);
    c[1] = (float)(cl & 1ull);
    c[2] = (float)__builtin_popcountll((cl >> LC_EDGE) & 0x3FFFull);
    c[3] = (float)__builtin_popcountll(cl >> LC_TRI);
    c[4] = (float)__builtin_popcount(ch);
    c[5] = eps_ch;
    c[6] = clampf((gv - mn) * inv_eps);
    c[7] = clampf((mx - gv) * inv_eps);
    uint16_t* o = feat + (size_t)v * CH;
    for (int k = 0; k < CH; k++) o[k] = f2bf(c[k]);
}

// MODE 0: Y = ReLU(conv(X, W) + b) as bf16; MODE 1: Y = conv(X, W) + b, channels 0..1 as fp32;
// MODE 2: dX(v) = ReLU'(M(v)) * sum_tap W[tap]^T dY(v - off(tap)) as bf16 (X is dY, M the
// layer's input, whose positive entries are those where the ReLU below passed its input).
// Waves stride over work items of CONV_YC consecutive 16-voxel runs along y (one x0, one z); each
// lane keeps its B fragments of all 27 taps in registers. Per item a wave stages the rows of X
// the item's runs need (y0 - 1..y0 + CONV_YC, z - 1..z + 1, over x0 - 1..x0 + 16, channel-last as
// in memory) in its own part of LDS with 16-byte loads, 3 (CONV_YC + 2) rows where one run at a
// time would stage 9 CONV_YC; each run then reads its taps' A fragments (4 channels of one voxel)
// from there, into three independent accumulators so that consecutive MFMAs do not wait on one
// another.
constexpr int CONV_YC = 4;
template <int MODE>
__global__ __launch_bounds__(BLOCK) void k_conv(const uint16_t* __restrict__ X, const uint16_t* __restrict__ W,
                                                const float* __restrict__ bias, const uint16_t* __restrict__ M,
                                                int nx, int ny, int nz, uint16_t* Yb, float* Yf) {
    constexpr int NR = CONV_YC + 2;  // staged rows per z plane
    __shared__ __attribute__((aligned(16))) uint16_t xt[BLOCK / 64][3][NR][TXC][CH];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, i = lane & 15, q = lane >> 4;
    const int nxs = (nx + 15) >> 4, nys = (ny + CONV_YC - 1) / CONV_YC;
    const long S = (long)nxs * nys * nz;
    const int sgn = MODE == 2 ? -1 : 1;
    const float bi = MODE == 2 ? 0.f : bias[i];
    short4_t bw[TAPS];
#pragma unroll
    for (int tap = 0; tap < TAPS; tap++)
        for (int r = 0; r < 4; r++)
            bw[tap][r] = (short)(MODE == 2 ? W[tap * CH * CH + i * CH + 4 * q + r]      // W[tap][ci = i][co = 4q + r]
                                           : W[tap * CH * CH + (4 * q + r) * CH + i]);  // W[tap][ci = 4q + r][co = i]
    // items of one wave are ordered by wave barriers only (waves of a block take different numbers)
    for (long w = (long)blockIdx.x * (BLOCK / 64) + wv; w < S; w += (long)gridDim.x * (BLOCK / 64)) {
        const int x0 = (int)(w % nxs) * 16;
        const long rest = w / nxs;
        const int y0 = (int)(rest % nys) * CONV_YC, z = (int)(rest / nys);
        for (int e = lane; e < 3 * NR * TXC * 2; e += 64) {  // 16-byte chunks (8 channels)
            const int row = e / (TXC * 2), rem = e % (TXC * 2), xi = rem >> 1, half = rem & 1;
            const int pz = row / NR, py = row % NR;
            const int yy = y0 - 1 + py, zz = z - 1 + pz, xx = x0 - 1 + xi;
            uint4 val = make_uint4(0u, 0u, 0u, 0u);
            if (xx >= 0 && xx < nx && yy >= 0 && yy < ny && zz >= 0 && zz < nz)
                val = *(const uint4*)(X + (((long)zz * ny + yy) * nx + xx) * CH + 8 * half);
            *(uint4*)&xt[wv][pz][py][xi][8 * half] = val;
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
        const int x = x0 + i;
        for (int ry = 0; ry < CONV_YC; ry++) {
            const int y = y0 + ry;
            if (y >= ny) break;  // uniform over the wave
            float4_t acc3[3] = {{0.f, 0.f, 0.f, 0.f}, {0.f, 0.f, 0.f, 0.f}, {0.f, 0.f, 0.f, 0.f}};
#pragma unroll
            for (int tap = 0; tap < TAPS; tap++) {
                int dx, dy, dz;
                tap_offset(tap, dx, dy, dz);
                short4_t a = {0, 0, 0, 0};
                if (x < nx) a = *(const short4_t*)&xt[wv][sgn * dz + 1][ry + sgn * dy + 1][i + sgn * dx + 1][4 * q];
                acc3[tap % 3] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, bw[tap], acc3[tap % 3], 0, 0, 0);
            }
            const float4_t acc = acc3[0] + acc3[1] + acc3[2];
            for (int r = 0; r < 4; r++) {
                const int vx = x0 + 4 * q + r;
                if (vx >= nx) continue;
                const long v = ((long)z * ny + y) * nx + vx;
                const float val = acc[r];
                if (MODE == 0) {
                    const float o = val + bi;
                    Yb[v * CH + i] = f2bf(o > 0.f ? o : 0.f);
                } else if (MODE == 1) {
                    if (i < COUT_LAST) Yf[v * COUT_LAST + i] = val + bi;
                } else {
                    Yb[v * CH + i] = bf2f(M[v * CH + i]) > 0.f ? f2bf(val) : (uint16_t)0;
                }
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
    }
}

// dW[tap][ci][co] += sum_v X(v + off(tap))[ci] * dY(v)[co], db[co] += sum_v dY(v)[co]. Each wave
// strides over 16-voxel runs. Per run it stages, in its own part of LDS and channel-major, the 9
// rows (y - 1..y + 1, z - 1..z + 1) of X over x0 - 1..x0 + 16 and the run's dY, loading 16-byte
// chunks; the 27 taps then read their A fragments (4 voxels of one channel) from LDS and
// accumulate into 27 fragments. The block sums its waves in LDS before one global atomic per
// entry. Runs of one wave are ordered by wave barriers only (waves of a block take different
// numbers of runs).
__global__ __launch_bounds__(BLOCK) void k_wgrad(const uint16_t* __restrict__ X, const uint16_t* __restrict__ dY,
                                                 int nx, int ny, int nz, float* dW, float* db) {
    __shared__ float wacc[WSZ];
    __shared__ float bacc[CH];
    __shared__ uint16_t xt[BLOCK / 64][9][CH][TXC];
    __shared__ __attribute__((aligned(16))) uint16_t yt[BLOCK / 64][CH][16];
    for (int k = threadIdx.x; k < WSZ; k += BLOCK) wacc[k] = 0.f;
    if (threadIdx.x < CH) bacc[threadIdx.x] = 0.f;
    __syncthreads();
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, i = lane & 15, q = lane >> 4;
    const int nxs = (nx + 15) >> 4;
    const long S = (long)nxs * ny * nz;
    float4_t acc[TAPS];  // fully unrolled loops keep the 27 fragments in registers
#pragma unroll
    for (int t = 0; t < TAPS; t++) acc[t] = (float4_t){0.f, 0.f, 0.f, 0.f};
    float bsum = 0.f;
    for (long w = (long)blockIdx.x * (BLOCK / 64) + wv; w < S; w += (long)gridDim.x * (BLOCK / 64)) {
        int x0, y, z;
        run_origin(w, nxs, ny, x0, y, z);
        for (int e = lane; e < 9 * TXC * 2; e += 64) {  // X rows: 16-byte chunks (8 channels)
            const int row = e / (TXC * 2), rem = e % (TXC * 2), xi = rem >> 1, half = rem & 1;
            const int yy = y + row % 3 - 1, zz = z + row / 3 - 1, xx = x0 - 1 + xi;
            uint4 val = make_uint4(0u, 0u, 0u, 0u);
            if (xx >= 0 && xx < nx && yy >= 0 && yy < ny && zz >= 0 && zz < nz)
                val = *(const uint4*)(X + (((long)zz * ny + yy) * nx + xx) * CH + 8 * half);
            const uint32_t wd[4] = {val.x, val.y, val.z, val.w};
            for (int c = 0; c < 8; c++) xt[wv][row][8 * half + c][xi] = (uint16_t)(wd[c >> 1] >> (16 * (c & 1)));
        }
        if (lane < 32) {  // dY of the run
            const int xi = lane >> 1, half = lane & 1, xx = x0 + xi;
            uint4 val = make_uint4(0u, 0u, 0u, 0u);
            if (xx < nx) val = *(const uint4*)(dY + (((long)z * ny + y) * nx + xx) * CH + 8 * half);
            const uint32_t wd[4] = {val.x, val.y, val.z, val.w};
            for (int c = 0; c < 8; c++) yt[wv][8 * half + c][xi] = (uint16_t)(wd[c >> 1] >> (16 * (c & 1)));
        }
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        const short4_t b = *(const short4_t*)&yt[wv][i][4 * q];  // B[k = 4q + r][co = i]
        for (int r = 0; r < 4; r++) bsum += bf2f((uint16_t)b[r]);
#pragma unroll
        for (int tap = 0; tap < TAPS; tap++) {
            const int dx = tap % 3 - 1, row = tap / 3;  // row = (dz + 1) * 3 + (dy + 1)
            short4_t a;
            for (int r = 0; r < 4; r++) a[r] = (short)xt[wv][row][i][4 * q + r + dx + 1];  // A[ci = i][k]
            acc[tap] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[tap], 0, 0, 0);
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
    }
#pragma unroll
    for (int tap = 0; tap < TAPS; tap++)
        for (int r = 0; r < 4; r++) atomicAdd(&wacc[tap * CH * CH + (4 * q + r) * CH + i], acc[tap][r]);
    atomicAdd(&bacc[i], bsum);
    __syncthreads();
    for (int k = threadIdx.x; k < WSZ; k += BLOCK) atomicAdd(&dW[k], wacc[k]);
    if (threadIdx.x < CH) atomicAdd(&db[threadIdx.x], bacc[threadIdx.x]);
}

// Loss terms (sums[0] += weighted BCE, sums[1] += y (size - t)^2) and the output gradient
// dY (bf16, channels 0 and 1), both scaled by 1 / n.
__global__ __launch_bounds__(BLOCK) void k_loss(const float* Z, const float* lab, long n, float pos_weight,
                                                float lambda, uint16_t* dY, float* sums) {
    __shared__ float red[2][BLOCK];
    const long v = (long)blockIdx.x * BLOCK + threadIdx.x;
    float bce = 0.f, mse = 0.f;
    if (v < n) {
        const float z = Z[v * 2], m = Z[v * 2 + 1], y = lab[v * 2], t = lab[v * 2 + 1];
        const float sp_pos = (z > 0.f ? 0.f : -z) + log1pf(expf(-fabsf(z)));  // -log sigmoid(z)
        const float sp_neg = sp_pos + z;                                       // -log(1 - sigmoid(z))
        bce = pos_weight * y * sp_pos + (1.f - y) * sp_neg;
        const float p = 1.f / (1.f + expf(-z));
        const float dz = pos_weight * y * (p - 1.f) + (1.f - y) * p;
        mse = y * (m - t) * (m - t);
        const float dm = 2.f * lambda * y * (m - t);
        const float inv = 1.f / (float)n;
        uint16_t* o = dY + v * CH;
        o[0] = f2bf(dz * inv);
        o[1] = f2bf(dm * inv);
        for (int k = 2; k < CH; k++) o[k] = 0;
    }
    red[0][threadIdx.x] = bce;
    red[1][threadIdx.x] = mse;
    __syncthreads();
    for (int s = BLOCK / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            red[0][threadIdx.x] += red[0][threadIdx.x + s];
            red[1][threadIdx.x] += red[1][threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        atomicAdd(&sums[0], red[0][0]);
        atomicAdd(&sums[1], red[1][0]);
    }
}

// Adam on the fp32 master parameters (masked entries stay zero), then the bf16 weight copies.
__global__ __launch_bounds__(BLOCK) void k_adam(float* p, const float* g, float* m, float* v, const float* mask,
                                                int n, float lr, float b1, float b2, float eps, float bc1, float bc2,
                                                uint16_t* wb) {
    const int k = blockIdx.x * BLOCK + threadIdx.x;
    if (k >= n) return;
    const float gk = g[k] * mask[k];
    const float mk = b1 * m[k] + (1.f - b1) * gk;
    const float vk = b2 * v[k] + (1.f - b2) * gk * gk;
    m[k] = mk;
    v[k] = vk;
    const float pk = (p[k] - lr * (mk / bc1) / (sqrtf(vk / bc2) + eps)) * mask[k];
    p[k] = pk;
    const int layer = k / LP, j = k % LP;
    if (j < WSZ) wb[layer * WSZ + j] = f2bf(pk);
}

// bf16 copies of the weights of all layers.
__global__ __launch_bounds__(BLOCK) void k_to_bf16(const float* p, int n, uint16_t* wb) {
    const int k = blockIdx.x * BLOCK + threadIdx.x;
    if (k >= n) return;
    const int layer = k / LP, j = k % LP;
    if (j < WSZ) wb[layer * WSZ + j] = f2bf(p[k]);
}

// Warm start: vertices with a positive edit logit move down by size * eps (size clamped to
// [0, 2]), never below the snap of f - eps.
// warm[v] records which vertices it moved.
__global__ __launch_bounds__(BLOCK) void k_warm(const float* f, float* g, const float* Z, long n, double eps,
                                                unsigned long long* edits, uint8_t* warm) {
    const long v = (long)blockIdx.x * BLOCK + threadIdx.x;
    bool ed = false;
    if (v < n && Z[v * 2] > 0.f) {
        float s = Z[v * 2 + 1];
        s = s < 0.f ? 0.f : (s > 2.f ? 2.f : s);
        const float lo = EditArith<float>::snap(f[v], eps);
        float c = (float)((double)g[v] - (double)s * eps);
        c = c < lo ? lo : c;
        c = c > g[v] ? g[v] : c;
        ed = !(c == g[v]);
        g[v] = c;
    }
    if (v < n) warm[v] = ed ? 1 : 0;
    const unsigned long long b = __ballot(ed ? 1 : 0);
    if (__lane_id() == 0 && b) atomicAdd(edits, (unsigned long long)__popcll(b));
}

// A warm edit moves a vertex by an amount that no step count of the edit stream expresses, so
// such vertices are reported with counter 6, whose value the stream carries explicitly
// (main_3d.cu:1502-1508); counters 6 and 7 (restored to f) already carry theirs.
__global__ __launch_bounds__(BLOCK) void k_warm_counter(const uint8_t* warm, long n, uint8_t* counter) {
    const long v = (long)blockIdx.x * BLOCK + threadIdx.x;
    if (v < n && warm[v] && counter[v] < 6) counter[v] = 6;
}

inline unsigned blocks_for(long n) { return (unsigned)((n + BLOCK - 1) / BLOCK); }

}  // namespace

struct dmtz_net {
    int device = 0;
    float* params = nullptr;  // fp32 master parameters
    float* grads = nullptr;
    float* m = nullptr;
    float* v = nullptr;
    float* mask = nullptr;
    uint16_t* wb = nullptr;  // bf16 weights, NL x WSZ
    float* sums = nullptr;   // loss terms
    unsigned long long* edits = nullptr;
    int64_t step = 0;
    // activations for up to cap vertices
    long cap = 0;
    uint16_t* act[NL - 1] = {nullptr, nullptr, nullptr};
    uint16_t* grad[2] = {nullptr, nullptr};
    uint16_t* feat = nullptr;
    float* z = nullptr;
    uint8_t* warm = nullptr;  // vertices moved by the last warm start
};

namespace {

#define NK(x)                                                                  \
    do {                                                                       \
        hipError_t e_ = (x);                                                   \
        if (e_ != hipSuccess) {                                                \
            dmtz_internal::set_error(ctx, hipGetErrorString(e_));              \
            return DMTZ_E_HIP;                                                 \
        }                                                                      \
    } while (0)

void free_acts(dmtz_net* net) {
    for (auto& p : net->act) (void)hipFree(p), p = nullptr;
    for (auto& p : net->grad) (void)hipFree(p), p = nullptr;
    (void)hipFree(net->feat);
    (void)hipFree(net->z);
    (void)hipFree(net->warm);
    net->feat = nullptr;
    net->z = nullptr;
    net->warm = nullptr;
    net->cap = 0;
}

int ensure_acts(dmtz_ctx* ctx, dmtz_net* net, long n) {
    if (n <= net->cap) return 0;
    free_acts(net);
    const size_t b = sizeof(uint16_t) * CH * (size_t)n;
    for (auto& p : net->act) NK(hipMalloc(&p, b));
    for (auto& p : net->grad) NK(hipMalloc(&p, b));
    NK(hipMalloc(&net->feat, b));
    NK(hipMalloc(&net->z, sizeof(float) * COUT_LAST * (size_t)n));
    NK(hipMalloc(&net->warm, (size_t)n));
    net->cap = n;
    return 0;
}

int check_grid(dmtz_ctx* ctx, int64_t nx, int64_t ny, int64_t nz) {
    if (nx < 3 || ny < 2 || nz < 2 || nx > (1 << 20) || ny > (1 << 20) || nz > (1 << 20) ||
        nx * ny * nz >= ((int64_t)1 << 31)) {
        dmtz_internal::set_error(ctx, "grid must have nx >= 3, ny >= 2, nz >= 2 and fewer than 2^31 vertices");
        return DMTZ_E_ARG;
    }
    return 0;
}

// Blocks of the convolution kernels: four waves each, about eight runs per wave, at most 4096.
unsigned conv_blocks(int64_t nx, int64_t ny, int64_t nz) {
    const long runs = ((nx + 15) / 16) * ny * nz;
    const long b = (runs + 31) / 32;
    return (unsigned)(b < 1 ? 1 : (b > 4096 ? 4096 : b));
}

// k_conv's grid: its work items are CONV_YC runs each, two items per wave.
unsigned conv_item_blocks(int64_t nx, int64_t ny, int64_t nz) {
    const long items = ((nx + 15) / 16) * ((ny + CONV_YC - 1) / CONV_YC) * nz;
    const long b = (items + 7) / 8;
    return (unsigned)(b < 1 ? 1 : (b > 4096 ? 4096 : b));
}

// Forward pass from feat; the output layer's result goes to out (2 floats per vertex).
int forward(dmtz_ctx* ctx, dmtz_net* net, const uint16_t* feat, int64_t nx, int64_t ny, int64_t nz, float* out) {
    hipStream_t s = dmtz_internal::stream(ctx);
    const unsigned gb = conv_item_blocks(nx, ny, nz);
    const uint16_t* in = feat;
    for (int l = 0; l < NL - 1; l++) {
        k_conv<0><<<gb, BLOCK, 0, s>>>(in, net->wb + l * WSZ, net->params + l * LP + WSZ, nullptr, (int)nx, (int)ny,
                                       (int)nz, net->act[l], nullptr);
        NK(hipGetLastError());
        in = net->act[l];
    }
    k_conv<1><<<gb, BLOCK, 0, s>>>(in, net->wb + (NL - 1) * WSZ, net->params + (NL - 1) * LP + WSZ, nullptr, (int)nx,
                                   (int)ny, (int)nz, nullptr, out);
    NK(hipGetLastError());
    return 0;
}

// Forward and backward pass: net->grads receives the parameter gradient, net->sums the loss
// terms.
int backprop(dmtz_ctx* ctx, dmtz_net* net, const uint16_t* feat, const float* labels, int64_t nx, int64_t ny,
             int64_t nz, float pos_weight, float lambda) {
    const long n = (long)(nx * ny * nz);
    int rc = ensure_acts(ctx, net, n);
    if (rc) return rc;
    rc = forward(ctx, net, feat, nx, ny, nz, net->z);
    if (rc) return rc;
    hipStream_t s = dmtz_internal::stream(ctx);
    NK(hipMemsetAsync(net->grads, 0, sizeof(float) * DMTZ_NET_PARAMS, s));
    NK(hipMemsetAsync(net->sums, 0, sizeof(float) * 2, s));
    k_loss<<<blocks_for(n), BLOCK, 0, s>>>(net->z, labels, n, pos_weight, lambda, net->grad[0], net->sums);
    NK(hipGetLastError());
    const unsigned gb = conv_item_blocks(nx, ny, nz);
    const long runs = ((nx + 15) / 16) * ny * nz;
    const unsigned wg = (unsigned)std::min<long>(512, (runs + 3) / 4);
    int cur = 0;  // grad[cur] holds the gradient at the current layer's output (before its ReLU)
    for (int l = NL - 1; l >= 0; l--) {
        const uint16_t* in = l == 0 ? feat : net->act[l - 1];
        k_wgrad<<<wg, BLOCK, 0, s>>>(in, net->grad[cur], (int)nx, (int)ny, (int)nz, net->grads + l * LP,
                                     net->grads + l * LP + WSZ);
        NK(hipGetLastError());
        if (l > 0) {
            k_conv<2><<<gb, BLOCK, 0, s>>>(net->grad[cur], net->wb + l * WSZ, nullptr, in, (int)nx, (int)ny, (int)nz,
                                           net->grad[1 - cur], nullptr);
            NK(hipGetLastError());
            cur = 1 - cur;
        }
    }
    return 0;
}

// Mean loss terms over the n vertices of the last backward pass.
int read_loss(dmtz_ctx* ctx, dmtz_net* net, long n, double* loss_out) {
    float h[2];
    hipStream_t s = dmtz_internal::stream(ctx);
    NK(hipMemcpyAsync(h, net->sums, sizeof(h), hipMemcpyDeviceToHost, s));
    NK(hipStreamSynchronize(s));
    if (loss_out) {
        loss_out[0] = h[0] / (double)n;
        loss_out[1] = h[1] / (double)n;
    }
    return 0;
}

}  // namespace

extern "C" {

int dmtz_net_create(dmtz_ctx* ctx, uint64_t seed, dmtz_net** out) {
    if (!ctx || !out) return DMTZ_E_ARG;
    *out = nullptr;
    dmtz_net* net = new dmtz_net();
    net->device = dmtz_internal::device(ctx);
    (void)hipSetDevice(net->device);
    std::vector<float> p(DMTZ_NET_PARAMS, 0.f), mask(DMTZ_NET_PARAMS, 0.f);
    std::mt19937_64 rng(seed);
    for (int l = 0; l < NL; l++) {
        const int cin = l == 0 ? CIN0 : CH, cout = l == NL - 1 ? COUT_LAST : CH;
        std::normal_distribution<float> nd(0.f, std::sqrt(2.f / (float)(TAPS * cin)));
        for (int t = 0; t < TAPS; t++)
            for (int ci = 0; ci < cin; ci++)
                for (int co = 0; co < cout; co++) {
                    const int k = l * LP + t * CH * CH + ci * CH + co;
                    p[k] = nd(rng);
                    mask[k] = 1.f;
                }
        for (int co = 0; co < cout; co++) mask[l * LP + WSZ + co] = 1.f;
    }
    auto fail = [&](hipError_t e) {
        dmtz_internal::set_error(ctx, hipGetErrorString(e));
        dmtz_net_destroy(net);
        return DMTZ_E_HIP;
    };
    hipError_t e = hipSuccess;
    for (float** b : {&net->params, &net->grads, &net->m, &net->v, &net->mask})
        if (e == hipSuccess) e = hipMalloc(b, sizeof(float) * DMTZ_NET_PARAMS);
    if (e == hipSuccess) e = hipMalloc(&net->wb, sizeof(uint16_t) * NL * WSZ);
    if (e == hipSuccess) e = hipMalloc(&net->sums, sizeof(float) * 2);
    if (e == hipSuccess) e = hipMalloc(&net->edits, sizeof(unsigned long long));
    if (e != hipSuccess) return fail(e);
    *out = net;
    int rc = dmtz_net_set_params(ctx, net, p.data());
    if (rc) {
        *out = nullptr;
        dmtz_net_destroy(net);
        return rc;
    }
    hipStream_t s = dmtz_internal::stream(ctx);
    e = hipMemcpyAsync(net->mask, mask.data(), sizeof(float) * DMTZ_NET_PARAMS, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
        *out = nullptr;
        return fail(e);
    }
    return 0;
}

void dmtz_net_destroy(dmtz_net* net) {
    if (!net) return;
    (void)hipSetDevice(net->device);
    free_acts(net);
    for (float* b : {net->params, net->grads, net->m, net->v, net->mask, net->sums}) (void)hipFree(b);
    (void)hipFree(net->wb);
    (void)hipFree(net->edits);
    delete net;
}

int dmtz_net_params(dmtz_ctx* ctx, dmtz_net* net, float* out) {
    if (!ctx || !net || !out) return DMTZ_E_ARG;
    hipStream_t s = dmtz_internal::stream(ctx);
    NK(hipMemcpyAsync(out, net->params, sizeof(float) * DMTZ_NET_PARAMS, hipMemcpyDeviceToHost, s));
    NK(hipStreamSynchronize(s));
    return 0;
}

int dmtz_net_set_params(dmtz_ctx* ctx, dmtz_net* net, const float* in) {
    if (!ctx || !net || !in) return DMTZ_E_ARG;
    hipStream_t s = dmtz_internal::stream(ctx);
    NK(hipMemcpyAsync(net->params, in, sizeof(float) * DMTZ_NET_PARAMS, hipMemcpyHostToDevice, s));
    NK(hipMemsetAsync(net->m, 0, sizeof(float) * DMTZ_NET_PARAMS, s));
    NK(hipMemsetAsync(net->v, 0, sizeof(float) * DMTZ_NET_PARAMS, s));
    k_to_bf16<<<blocks_for(DMTZ_NET_PARAMS), BLOCK, 0, s>>>(net->params, DMTZ_NET_PARAMS, net->wb);
    NK(hipGetLastError());
    NK(hipStreamSynchronize(s));
    net->step = 0;
    return 0;
}

int dmtz_net_features_f32(dmtz_ctx* ctx, const float* d_g, int64_t nx, int64_t ny, int64_t nz, double eps,
                          double range, uint16_t* d_feat) {
    if (!ctx || !d_g || !d_feat || !(eps > 0) || !(range > 0)) return DMTZ_E_ARG;
    int rc = check_grid(ctx, nx, ny, nz);
    if (rc) return rc;
    Dims d(nx, ny, nz);
    const i64 n = nx * ny * nz;
    hipStream_t s = dmtz_internal::stream(ctx);
    k_features<<<blocks_for(n), BLOCK, 0, s>>>(d_g, d, n, (float)(1.0 / eps), (float)(std::log10(eps / range) + 3.0),
                                                d_feat);
    NK(hipGetLastError());
    NK(hipStreamSynchronize(s));
    return 0;
}

int dmtz_net_forward(dmtz_ctx* ctx, dmtz_net* net, const uint16_t* d_feat, int64_t nx, int64_t ny, int64_t nz,
                     float* d_out) {
    if (!ctx || !net || !d_feat || !d_out) return DMTZ_E_ARG;
    int rc = check_grid(ctx, nx, ny, nz);
    if (rc) return rc;
    rc = ensure_acts(ctx, net, (long)(nx * ny * nz));
    if (rc) return rc;
    rc = forward(ctx, net, d_feat, nx, ny, nz, d_out);
    if (rc) return rc;
    NK(hipStreamSynchronize(dmtz_internal::stream(ctx)));
    return 0;
}

int dmtz_net_gradient(dmtz_ctx* ctx, dmtz_net* net, const uint16_t* d_feat, const float* d_labels, int64_t nx,
                      int64_t ny, int64_t nz, float pos_weight, float lambda, float* grads_out, double* loss_out) {
    if (!ctx || !net || !d_feat || !d_labels) return DMTZ_E_ARG;
    int rc = check_grid(ctx, nx, ny, nz);
    if (rc) return rc;
    rc = backprop(ctx, net, d_feat, d_labels, nx, ny, nz, pos_weight, lambda);
    if (rc) return rc;
    hipStream_t s = dmtz_internal::stream(ctx);
    if (grads_out) NK(hipMemcpyAsync(grads_out, net->grads, sizeof(float) * DMTZ_NET_PARAMS, hipMemcpyDeviceToHost, s));
    return read_loss(ctx, net, (long)(nx * ny * nz), loss_out);
}

int dmtz_net_train_step(dmtz_ctx* ctx, dmtz_net* net, const uint16_t* d_feat, const float* d_labels, int64_t nx,
                        int64_t ny, int64_t nz, float pos_weight, float lambda, float lr, double* loss_out) {
    if (!ctx || !net || !d_feat || !d_labels || !(lr > 0)) return DMTZ_E_ARG;
    int rc = check_grid(ctx, nx, ny, nz);
    if (rc) return rc;
    rc = backprop(ctx, net, d_feat, d_labels, nx, ny, nz, pos_weight, lambda);
    if (rc) return rc;
    net->step++;
    const float b1 = 0.9f, b2 = 0.999f;
    const float bc1 = 1.f - std::pow(b1, (float)net->step), bc2 = 1.f - std::pow(b2, (float)net->step);
    hipStream_t s = dmtz_internal::stream(ctx);
    k_adam<<<blocks_for(DMTZ_NET_PARAMS), BLOCK, 0, s>>>(net->params, net->grads, net->m, net->v, net->mask,
                                                         DMTZ_NET_PARAMS, lr, b1, b2, 1e-8f, bc1, bc2, net->wb);
    NK(hipGetLastError());
    return read_loss(ctx, net, (long)(nx * ny * nz), loss_out);
}

int dmtz_correct_batch_learned_f32(dmtz_ctx* ctx, dmtz_net* net, const dmtz_field_f32* fields, int32_t nfields,
                                   const dmtz_flags* flags, dmtz_counts* counts_out, int32_t* status_out,
                                   int32_t* sweeps_out, int64_t* warm_edits_out) {
    if (!ctx || !net || !flags || nfields < 0 || (nfields > 0 && !fields)) return DMTZ_E_ARG;
    for (int32_t i = 0; i < nfields; i++) {
        const dmtz_field_f32& fd = fields[i];
        if (!fd.on_device || !fd.f || !fd.g) {
            dmtz_internal::set_error(ctx, "batched learned correction takes device fields only");
            return DMTZ_E_ARG;
        }
        int rc = check_grid(ctx, fd.nx, fd.ny, fd.nz);
        if (rc) return rc;
        const long n = (long)(fd.nx * fd.ny * fd.nz);
        rc = ensure_acts(ctx, net, n);
        if (rc) return rc;
        rc = dmtz_net_features_f32(ctx, fd.g, fd.nx, fd.ny, fd.nz, fd.eps, fd.range, net->feat);
        if (rc) return rc;
        rc = forward(ctx, net, net->feat, fd.nx, fd.ny, fd.nz, net->z);
        if (rc) return rc;
        hipStream_t s = dmtz_internal::stream(ctx);
        NK(hipMemsetAsync(net->edits, 0, sizeof(unsigned long long), s));
        k_warm<<<blocks_for(n), BLOCK, 0, s>>>(fd.f, fd.g, net->z, n, fd.eps, net->edits, net->warm);
        NK(hipGetLastError());
        unsigned long long ne = 0;
        NK(hipMemcpyAsync(&ne, net->edits, sizeof(ne), hipMemcpyDeviceToHost, s));
        NK(hipStreamSynchronize(s));
        if (warm_edits_out) warm_edits_out[i] = (int64_t)ne;
    }
    return dmtz_correct_batch_f32(ctx, fields, nfields, flags, counts_out, status_out, sweeps_out);
}

int dmtz_correct_learned_f32(dmtz_ctx* ctx, dmtz_net* net, const float* d_f, float* d_g, int64_t nx, int64_t ny,
                             int64_t nz, double eps, double range, const dmtz_flags* flags, dmtz_counts* out,
                             uint8_t* d_counter_out, int32_t* sweeps_out, int64_t* warm_edits_out) {
    if (!ctx || !net || !d_f || !d_g || !flags) return DMTZ_E_ARG;
    int rc = check_grid(ctx, nx, ny, nz);
    if (rc) return rc;
    const long n = (long)(nx * ny * nz);
    rc = ensure_acts(ctx, net, n);
    if (rc) return rc;
    rc = dmtz_net_features_f32(ctx, d_g, nx, ny, nz, eps, range, net->feat);
    if (rc) return rc;
    rc = forward(ctx, net, net->feat, nx, ny, nz, net->z);
    if (rc) return rc;
    hipStream_t s = dmtz_internal::stream(ctx);
    NK(hipMemsetAsync(net->edits, 0, sizeof(unsigned long long), s));
    k_warm<<<blocks_for(n), BLOCK, 0, s>>>(d_f, d_g, net->z, n, eps, net->edits, net->warm);
    NK(hipGetLastError());
    unsigned long long ne = 0;
    NK(hipMemcpyAsync(&ne, net->edits, sizeof(ne), hipMemcpyDeviceToHost, s));
    NK(hipStreamSynchronize(s));
    if (warm_edits_out) *warm_edits_out = (int64_t)ne;
    const int st = dmtz_correct_dev_f32(ctx, d_f, d_g, nx, ny, nz, eps, range, flags, out, d_counter_out, sweeps_out);
    if (st < 0 || !d_counter_out) return st;
    k_warm_counter<<<blocks_for(n), BLOCK, 0, s>>>(net->warm, n, d_counter_out);
    NK(hipGetLastError());
    NK(hipStreamSynchronize(s));
    return st;
}

}  // extern "C"
