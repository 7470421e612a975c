// Context management and the type-independent entry points of include/dmtz.h.
#include <cstdio>

#include "driver.h"

namespace dmtz_internal {
int noalloc_setup(int device);  // dmtz_noalloc.hip: device address of its scratch array
}

extern "C" {

int dmtz_ctx_create(dmtz_ctx** out, int device) {
    if (!out) return DMTZ_E_ARG;
    *out = nullptr;
    dmtz_ctx* ctx = new dmtz_ctx();
    ctx->device = device;
    hipError_t e = hipSetDevice(device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipMalloc(&ctx->d_ctr, sizeof(ull) * K_NCTR);
    if (e == hipSuccess) e = hipHostMalloc(&ctx->h_ctr, sizeof(ull) * K_NCTR);
    if (e == hipSuccess) e = hipHostMalloc(&ctx->h_as_total, 2 * sizeof(ull));
    if (e == hipSuccess) e = hipEventCreate(&ctx->ev0);
    if (e == hipSuccess) e = hipEventCreate(&ctx->ev1);
    if (e == hipSuccess && dmtz_internal::noalloc_setup(device) != 0) e = hipErrorInvalidSymbol;
    if (e != hipSuccess) {
        std::fprintf(stderr, "dmtz_ctx_create: %s\n", hipGetErrorString(e));
        delete ctx;
        return DMTZ_E_HIP;
    }
    *out = ctx;
    return 0;
}

void dmtz_ctx_destroy(dmtz_ctx* ctx) {
    if (!ctx) return;
    for (dmtz_ctx* w : ctx->workers) dmtz_ctx_destroy(w);
    ctx->workers.clear();
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    if (ctx->as_stream) (void)hipStreamSynchronize(ctx->as_stream);
    free_bufs(ctx);
    free_rec(ctx);
    free_fc(ctx);
    free_paths(ctx);
    (void)hipFree(ctx->d_ctr);
    for (int k = 0; k < 3; k++) {
        (void)hipFree(ctx->d_fk[k]);
        (void)hipFree(ctx->d_fsorted[k]);
        (void)hipFree(ctx->d_ftmp[k]);
    }
    if (ctx->fix_stream) {
        (void)hipStreamSynchronize(ctx->fix_stream);
        (void)hipStreamDestroy(ctx->fix_stream);
    }
    if (ctx->fix_fork) (void)hipEventDestroy(ctx->fix_fork);
    if (ctx->fix_join) (void)hipEventDestroy(ctx->fix_join);
    if (ctx->as_stream) {
        (void)hipStreamSynchronize(ctx->as_stream);
        (void)hipStreamDestroy(ctx->as_stream);
    }
    for (hipEvent_t ev : {ctx->as_ev_fork, ctx->as_ev_count, ctx->as_ev_done})
        if (ev) (void)hipEventDestroy(ev);
    (void)hipHostFree(ctx->h_ctr);
    (void)hipHostFree(ctx->h_as_total);
    if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
    if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

const char* dmtz_last_error(const dmtz_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int dmtz_eps_from_rel_f32(const float* f, int64_t n, double rel, double* e, double* r) {
    return eps_from_rel(f, n, rel, e, r);
}

int dmtz_eps_from_rel_f64(const double* f, int64_t n, double rel, double* e, double* r) {
    return eps_from_rel(f, n, rel, e, r);
}

int64_t dmtz_trajectory(const dmtz_ctx* ctx, int64_t* out, int64_t cap) {
    if (!ctx) return 0;
    int64_t n = (int64_t)ctx->traj.size() / 8;
    if (out)
        for (int64_t i = 0; i < n && i < cap; i++)
            for (int k = 0; k < 8; k++) out[i * 8 + k] = ctx->traj[(size_t)i * 8 + k];
    return n;
}

int dmtz_phase_times(const dmtz_ctx* ctx, double* out6) {
    if (!ctx || !out6) return DMTZ_E_ARG;
    for (int k = 0; k < 6; k++) out6[k] = ctx->phase[k];
    return 0;
}

int dmtz_kernel_stats(dmtz_ctx* ctx, double* out4, int reset) {
    if (!ctx || !out4) return DMTZ_E_ARG;
    out4[0] = ctx->det_launches;
    out4[1] = ctx->det_ms;
    out4[2] = ctx->det_bytes;
    out4[3] = ctx->det_voxels;
    if (reset) ctx->det_launches = ctx->det_ms = ctx->det_bytes = ctx->det_voxels = 0;
    return 0;
}

int dmtz_run_info(const dmtz_ctx* ctx, int64_t* out4) {
    if (!ctx || !out4) return DMTZ_E_ARG;
    out4[0] = ctx->stalls;
    out4[1] = ctx->finished;
    out4[2] = ctx->radius;
    out4[3] = (int64_t)ctx->dirty_total;
    return 0;
}

}  // extern "C"

// Context internals for the library's other translation unit (learned.hip).
namespace dmtz_internal {
hipStream_t stream(dmtz_ctx* ctx) { return ctx->stream; }
int device(dmtz_ctx* ctx) { return ctx->device; }
void set_error(dmtz_ctx* ctx, const char* msg) { ctx->err = msg; }
}  // namespace dmtz_internal


// Field files (getdata2 main_3d.cu:153-181, saveVectorToBin main_3d.cu:1315-1327). Host code.
namespace {
int read_doubles(FILE* fp, int64_t count, double* out, int64_t cap) {
    if (!out) return 0;
    if (cap < count) return DMTZ_E_CAPACITY;
    return (int64_t)std::fread(out, sizeof(double), (size_t)count, fp) == count ? 0 : DMTZ_E_IO;
}
}  // namespace

extern "C" {

int dmtz_read_raw_f64(const char* path, double* out, int64_t cap, int64_t* n_out) {
    if (!path || !n_out) return DMTZ_E_ARG;
    FILE* fp = std::fopen(path, "rb");
    if (!fp) return DMTZ_E_IO;
    int rc = 0;
    long size = -1;
    if (std::fseek(fp, 0, SEEK_END) == 0) size = std::ftell(fp);
    if (size < 0 || std::fseek(fp, 0, SEEK_SET) != 0) rc = DMTZ_E_IO;
    if (rc == 0) {
        *n_out = (int64_t)size / (int64_t)sizeof(double);  // a trailing partial value is ignored, as there
        rc = read_doubles(fp, *n_out, out, cap);
    }
    std::fclose(fp);
    return rc;
}

int dmtz_write_bin_f64(const char* path, const double* data, int64_t n) {
    if (!path || n < 0 || (n > 0 && !data)) return DMTZ_E_ARG;
    FILE* fp = std::fopen(path, "wb");
    if (!fp) return DMTZ_E_IO;
    const uint64_t count = (uint64_t)n;
    bool ok = std::fwrite(&count, sizeof(count), 1, fp) == 1;
    ok = ok && (n == 0 || (int64_t)std::fwrite(data, sizeof(double), (size_t)n, fp) == n);
    ok = (std::fclose(fp) == 0) && ok;
    return ok ? 0 : DMTZ_E_IO;
}

int dmtz_read_bin_f64(const char* path, double* out, int64_t cap, int64_t* n_out) {
    if (!path || !n_out) return DMTZ_E_ARG;
    FILE* fp = std::fopen(path, "rb");
    if (!fp) return DMTZ_E_IO;
    uint64_t count = 0;
    int rc = std::fread(&count, sizeof(count), 1, fp) == 1 ? 0 : DMTZ_E_IO;
    long size = -1;
    if (rc == 0 && std::fseek(fp, 0, SEEK_END) == 0) size = std::ftell(fp);
    if (rc == 0 && (size < 0 || (uint64_t)(size - (long)sizeof(count)) != count * sizeof(double) ||
                    std::fseek(fp, (long)sizeof(count), SEEK_SET) != 0))
        rc = DMTZ_E_IO;  // the count must match the file's length
    if (rc == 0) {
        *n_out = (int64_t)count;
        rc = read_doubles(fp, *n_out, out, cap);
    }
    std::fclose(fp);
    return rc;
}

}  // extern "C"
