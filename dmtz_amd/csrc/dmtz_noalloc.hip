// dmtz_classify_noalloc_f32 (include/dmtz.h): the classifier on device-resident data without any
// allocation, for callers that cannot allocate device memory.
//
// Its scratch is a file-scope device array: the star codes of up to NOALLOC_MAX_VOXELS vertices
// followed by the four per-type counters. In host code the array's name is only a stand-in for
// the symbol, so its device address is looked up once per device with hipGetSymbolAddress, when
// the library's init call (dmtz_ctx_create) runs with that device current, and kept per device;
// every kernel takes that address as an argument. Each GPU keeps its own copy of the array until
// the process ends, and every call on a GPU shares it: the counters are cleared by a kernel at the
// start of each call, and calls that overlap on one GPU race on the array.
#include <mutex>

#include "driver.h"

namespace {

constexpr int64_t NOALLOC_MAX_VOXELS = (int64_t)1 << 21;  // 128^3 vertices
constexpr size_t NOALLOC_CODE_BYTES = sizeof(StarCode) * (size_t)NOALLOC_MAX_VOXELS;  // 64 MiB
constexpr size_t NOALLOC_BYTES = NOALLOC_CODE_BYTES + 256;                            // + counters

__device__ __align__(256) unsigned char scratch[NOALLOC_BYTES];

constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex noalloc_mu;
void* noalloc_ptr[NOALLOC_MAX_DEVICES] = {};  // device address of scratch, per device

__global__ __launch_bounds__(64) void k_clear_counters(ull* ctr) {
    if (threadIdx.x < 4) ctr[threadIdx.x] = 0;
}

// Critical cells each vertex owns, split into the two output masks of dmtz_classify_f32.
__global__ __launch_bounds__(BLOCK) void k_crit_split(const StarCode* code, i64 n, uint64_t* lo_out,
                                                      uint32_t* hi_out) {
    const i64 v = (i64)blockIdx.x * BLOCK + threadIdx.x;
    if (v >= n) return;
    uint64_t lo, ilo;
    uint32_t hi, ihi;
    code_masks(code[v], lo, hi, ilo, ihi);
    if (lo_out) lo_out[v] = lo & ilo;
    if (hi_out) hi_out[v] = hi & ihi;
}

}  // namespace

namespace dmtz_internal {
// Called by dmtz_ctx_create with `device` current.
int noalloc_setup(int device) {
    if (device < 0 || device >= NOALLOC_MAX_DEVICES) return DMTZ_E_ARG;
    std::lock_guard<std::mutex> lk(noalloc_mu);
    if (noalloc_ptr[device]) return 0;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) return DMTZ_E_HIP;
    noalloc_ptr[device] = p;
    return 0;
}
}  // namespace dmtz_internal

extern "C" {

int dmtz_classify_noalloc_f32(dmtz_ctx* ctx, const float* d_field, int64_t nx, int64_t ny, int64_t nz,
                              int64_t* counts4, uint64_t* d_crit_lo, uint32_t* d_crit_hi) {
    int rc = check_dims(ctx, nx, ny, nz);
    if (rc) return rc;
    if (!d_field || !counts4) {
        ctx->err = "null argument";
        return DMTZ_E_ARG;
    }
    const i64 N = nx * ny * nz;
    if (N > NOALLOC_MAX_VOXELS) {
        ctx->err = "dmtz_classify_noalloc_f32 takes at most 2^21 vertices";
        return DMTZ_E_CAPACITY;
    }
    CK(hipSetDevice(ctx->device));
    void* base = nullptr;
    {
        std::lock_guard<std::mutex> lk(noalloc_mu);
        if (ctx->device >= 0 && ctx->device < NOALLOC_MAX_DEVICES) base = noalloc_ptr[ctx->device];
    }
    if (!base) {
        ctx->err = "scratch of dmtz_classify_noalloc_f32 is not set up on this device";
        return DMTZ_E_HIP;
    }
    StarCode* code = (StarCode*)base;
    ull* ctr = (ull*)((unsigned char*)base + NOALLOC_CODE_BYTES);
    Dims d{nx, ny, nz};
    Counts c = make_counts(d);
    hipStream_t s = ctx->stream;
    k_clear_counters<<<1, 64, 0, s>>>(ctr);
    CK(hipGetLastError());
    k_code<float><<<nblocks(N), BLOCK, 0, s>>>(d_field, d, c, code, nullptr);
    CK(hipGetLastError());
    k_crit_count<<<nblocks(N), BLOCK, 0, s>>>(code, N, ctr);
    CK(hipGetLastError());
    if (d_crit_lo || d_crit_hi) {
        k_crit_split<<<nblocks(N), BLOCK, 0, s>>>(code, N, d_crit_lo, d_crit_hi);
        CK(hipGetLastError());
    }
    // the counters go through the context's pinned host buffer (allocated by dmtz_ctx_create)
    CK(hipMemcpyAsync(ctx->h_ctr, ctr, sizeof(ull) * 4, hipMemcpyDeviceToHost, s));
    CK(hipStreamSynchronize(s));
    for (int k = 0; k < 4; k++) counts4[k] = (int64_t)ctx->h_ctr[k];
    return 0;
}

}  // extern "C"
