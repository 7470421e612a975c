// Entry points of include/dmtz.h for float32 fields.
#include "driver.h"

extern "C" {

int dmtz_correct_f32(dmtz_ctx* ctx, const float* f, float* g, int64_t nx, int64_t ny, int64_t nz, double eps,
                     double range, const dmtz_flags* fl, dmtz_counts* out, uint8_t* counter_out, int32_t* sweeps) {
    if (!ctx) return DMTZ_E_ARG;
    (void)hipSetDevice(ctx->device);
    return correct_host<float>(ctx, f, g, nx, ny, nz, eps, range, fl, out, counter_out, sweeps);
}

int dmtz_correct_dev_f32(dmtz_ctx* ctx, const float* f, float* g, int64_t nx, int64_t ny, int64_t nz, double eps,
                         double range, const dmtz_flags* fl, dmtz_counts* out, uint8_t* counter_out, int32_t* sweeps) {
    if (!ctx) return DMTZ_E_ARG;
    (void)hipSetDevice(ctx->device);
    return correct_dev<float>(ctx, f, g, nx, ny, nz, eps, range, fl, out, counter_out, sweeps);
}

int dmtz_correct_batch_f32(dmtz_ctx* ctx, const dmtz_field_f32* fields, int32_t nfields, const dmtz_flags* flags,
                           dmtz_counts* counts_out, int32_t* status_out, int32_t* sweeps_out) {
    if (!ctx) return DMTZ_E_ARG;
    (void)hipSetDevice(ctx->device);
    return correct_batch<float>(ctx, fields, nfields, flags, counts_out, status_out, sweeps_out);
}

int dmtz_classify_f32(dmtz_ctx* ctx, const float* field, int64_t nx, int64_t ny, int64_t nz, int64_t* counts4,
                      uint64_t* crit_lo, uint32_t* crit_hi) {
    if (!ctx) return DMTZ_E_ARG;
    (void)hipSetDevice(ctx->device);
    return classify_host<float>(ctx, field, nx, ny, nz, counts4, crit_lo, crit_hi);
}

int dmtz_gradient_f32(dmtz_ctx* ctx, const float* fld, int64_t nx, int64_t ny, int64_t nz, int64_t* pv, int64_t* pe,
                      int64_t* pf, int64_t* pt) {
    if (!ctx) return DMTZ_E_ARG;
    (void)hipSetDevice(ctx->device);
    return gradient_host<float>(ctx, fld, nx, ny, nz, pv, pe, pf, pt);
}

int dmtz_targets_f32(dmtz_ctx* ctx, const float* f, const float* g, int64_t nx, int64_t ny, int64_t nz, double eps,
                     double range, const dmtz_flags* fl, uint8_t* targets, int64_t* counts4) {
    if (!ctx) return DMTZ_E_ARG;
    (void)hipSetDevice(ctx->device);
    return targets_host<float>(ctx, f, g, nx, ny, nz, eps, range, fl, targets, counts4);
}

int dmtz_paths_f32(dmtz_ctx* ctx, const float* f, const float* g, int64_t nx, int64_t ny, int64_t nz,
                   const dmtz_flags* fl, int64_t* counts2, int64_t* ds_id, int64_t* ds_rec, int64_t* as_id,
                   int64_t* as_rec, int64_t cap, uint8_t* targets) {
    if (!ctx) return DMTZ_E_ARG;
    (void)hipSetDevice(ctx->device);
    return paths_host<float>(ctx, f, g, nx, ny, nz, fl, counts2, ds_id, ds_rec, as_id, as_rec, cap, targets);
}

int dmtz_walls_f32(dmtz_ctx* ctx, const float* f, const float* g, int64_t nx, int64_t ny, int64_t nz, double eps,
                   double range, const dmtz_flags* flags, int64_t* count, int64_t* face, int64_t* edge, int64_t cap,
                   uint8_t* targets) {
    if (!ctx) return DMTZ_E_ARG;
    (void)hipSetDevice(ctx->device);
    return walls_host<float>(ctx, f, g, nx, ny, nz, eps, range, flags, count, face, edge, cap, targets);
}

}  // extern "C"
