/* dmtz.h — C ABI of the MI355X topology-correction path.
 *
 * The reference (YuxiaoLi1234/DMTz, main_3d.cu) has no library interface: its hot path is the
 * device-resident span of main() from the host-to-device copy of f and g (main_3d.cu:6384-6393)
 * to the device-to-host copy of the corrected field (main_3d.cu:7156), the span it times as
 * additional_time (main_3d.cu:6384, 7152-7155). Its inputs arrive through argv and raw files
 * (main_3d.cu:6260-6271, 153-181, 6300, 6364) and its state lives in __device__ globals
 * (main_3d.cu:43-129). Each entry point below replaces one piece of that span; the comment on
 * each names the reference lines it stands in for.
 *
 * Conventions
 *   - Fields are contiguous, x fastest: index = x + y*nx + z*nx*ny (main_3d.cu:1945-1947).
 *   - Grids need nx >= 3, ny >= 2, nz >= 2 and nx*ny*nz < 2^31; values must be finite.
 *   - Return codes: 0 = converged / OK, 1 = not converged within max_sweeps, 2 = stalled,
 *     negative = error (see DMTZ_E_*); dmtz_last_error() gives a message. Nothing calls exit().
 *   - A context owns one HIP stream and its scratch memory; calls on one context are
 *     host-synchronous and a context must not be used from two threads at once.
 */
#ifndef DMTZ_H
#define DMTZ_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DMTZ_OK 0
#define DMTZ_NOT_CONVERGED 1
/* The loop reached a fixed point with mismatches left: a sweep changed no value and no counter
 * and its counts repeat, so every later sweep repeats it (the reference loops forever there:
 * main_3d.cu:6784 has no exit). The returned field is the state the reference keeps. */
#define DMTZ_STALLED 2
#define DMTZ_E_ARG (-1)
#define DMTZ_E_HIP (-2)
#define DMTZ_E_CAPACITY (-3)
#define DMTZ_E_UNSUPPORTED (-4)
#define DMTZ_E_WALK (-5)
/* Batch status of a field that never ran because an earlier field failed with an error. */
#define DMTZ_E_NOT_RUN (-6)
#define DMTZ_E_IO (-7)

/* The reference's positional flags (main_3d.cu:6264-6271) plus a sweep cap and a finisher.
 * max_sweeps <= 0 selects the default cap (100000). mode: 0 = plain loop.
 * finish = 1 (no counterpart in the reference): on the k-th stall, snap every vertex within
 * Chebyshev distance 2^(k-1) of the edit sets of the stalled sweeps (g = the snap of f - eps,
 * counter 6: the terminal edit of main_3d.cu:1140-1141, which the reference never undoes). When those boxes
 * would cover the grid, or the edit sets are empty, restore g = f on every vertex instead
 * (counter 7, never edited again). Edits only lower g or raise counters and the radius doubles
 * with every stall, so the loop then ends with zero mismatches and |g - f| <= eps. Runs that
 * never stall are unchanged.
 * preserve_connectors = 1 (with preserve_vpath = 1): the walls check of main_3d.cu:5064-5133
 * with a stack and visited set per start face instead of the reference's shared stack (its
 * defects make the reference's result depend on thread timing); see dmtz_walls_f32. */
typedef struct dmtz_flags {
    int32_t preserve_saddles;    /* argv[4]: fix 1- and 2-saddle mismatches            */
    int32_t preserve_vpath;      /* argv[5]: also match separatrices (V-paths)          */
    int32_t preserve_geometry;   /* argv[6]: 1 = path geometry, 0 = endpoint labels      */
    int32_t filtration;          /* argv[7]: pin f-critical cells to f - eps first       */
    int32_t preserve_connectors; /* argv[8]: also match 2-saddle walls (deterministic)   */
    int32_t preserve_type;       /* argv[9]: nonzero = every edit snaps to f - eps       */
    int32_t edit_type;           /* argv[10]: 0 = halve the gap, 1 = -2^-k eps steps     */
    int32_t max_sweeps;          /* cap on correction sweeps (the reference has none)    */
    int32_t mode;                /* 0 = plain                                           */
    int32_t finish;              /* 1 = snap boxes around stalled edit sets (see above)  */
} dmtz_flags;

/* Mismatch counters count_f_* (main_3d.cu:68-75) in the order min, 1-saddle, 2-saddle, max,
 * descending path, ascending path, descending wall, ascending wall. As in the reference
 * (main_3d.cu:6767-6770), ds_vpath, as_vpath and ds_wall start at 1 and keep that value until the
 * first separatrix check runs, which happens only once the critical-cell counts reach zero: a
 * run with preserve_vpath = 1 that stops earlier (max_sweeps or a stall) reports 1 there,
 * meaning "not checked yet", not a measured count. Runs with preserve_vpath = 0 report 0. */
typedef struct dmtz_counts {
    int64_t min, saddle1, saddle2, max, ds_vpath, as_vpath, ds_wall, as_wall;
} dmtz_counts;

typedef struct dmtz_ctx dmtz_ctx;

/* Context with its own stream on HIP device `device`. Replaces the global device state of
 * main_3d.cu:43-129 and the allocations of main_3d.cu:6465-6714. */
int dmtz_ctx_create(dmtz_ctx** out, int device);
void dmtz_ctx_destroy(dmtz_ctx* ctx);
const char* dmtz_last_error(const dmtz_ctx* ctx);

/* eps = rel * (max f - min f) and range = max f - min f, main_3d.cu:6301-6309. */
int dmtz_eps_from_rel_f32(const float* f, int64_t n, double rel, double* eps_out, double* range_out);
int dmtz_eps_from_rel_f64(const double* f, int64_t n, double rel, double* eps_out, double* range_out);

/* Full correction of g (host pointers): main_3d.cu:6384-7156 (H2D copy, gradient of f and g,
 * mismatch detection, the sweep loop of main_3d.cu:6784-7149, D2H copy of the corrected field).
 * g_inout is corrected in place. counter_out (nullable, nx*ny*nz bytes) receives delta_counter
 * (main_3d.cu:84, values 1..6; 7 = restored to f by the finisher). sweeps_out (nullable) receives the
 * number of sweeps. */
int dmtz_correct_f32(dmtz_ctx* ctx, const float* f, float* g_inout, int64_t nx, int64_t ny, int64_t nz,
                     double eps, double range, const dmtz_flags* flags, dmtz_counts* out,
                     uint8_t* counter_out, int32_t* sweeps_out);
int dmtz_correct_f64(dmtz_ctx* ctx, const double* f, double* g_inout, int64_t nx, int64_t ny, int64_t nz,
                     double eps, double range, const dmtz_flags* flags, dmtz_counts* out,
                     uint8_t* counter_out, int32_t* sweeps_out);

/* Same loop on device-resident fields (the span main_3d.cu:6394-7155, after the H2D copy):
 * d_f, d_g are device pointers on the context's device; d_counter_out is a nullable device
 * pointer. No host copies of field data are made. */
int dmtz_correct_dev_f32(dmtz_ctx* ctx, const float* d_f, float* d_g, int64_t nx, int64_t ny, int64_t nz,
                         double eps, double range, const dmtz_flags* flags, dmtz_counts* out,
                         uint8_t* d_counter_out, int32_t* sweeps_out);
int dmtz_correct_dev_f64(dmtz_ctx* ctx, const double* d_f, double* d_g, int64_t nx, int64_t ny, int64_t nz,
                         double eps, double range, const dmtz_flags* flags, dmtz_counts* out,
                         uint8_t* d_counter_out, int32_t* sweeps_out);

/* One field of a batch: f and g (corrected in place) in host memory, or in device memory on the
 * context's device when on_device != 0; eps and range as for dmtz_correct_*. */
typedef struct dmtz_field_f32 {
    const float* f;
    float* g;
    int64_t nx, ny, nz;
    double eps, range;
    int32_t on_device;
    int32_t reserved;
} dmtz_field_f32;
typedef struct dmtz_field_f64 {
    const double* f;
    double* g;
    int64_t nx, ny, nz;
    double eps, range;
    int32_t on_device;
    int32_t reserved;
} dmtz_field_f64;

/* Batch correction: every field of the array is corrected with the same flags (the reference runs
 * one field per process, main_3d.cu:6255). The fields run concurrently, largest first, in up to
 * 8 lanes: host threads with worker contexts of their own (stream and buffers, kept by ctx for
 * later batches), so that their sweeps overlap on the device. Each field's result is exactly
 * that of a single correction. counts_out, status_out and sweeps_out (each nullable) receive one
 * entry per field; status_out[i] is field i's return code, or DMTZ_E_NOT_RUN for a field that
 * never started because another field failed with an error (lanes stop taking fields then).
 * counts_out and sweeps_out are written only for fields that ran to a non-negative status; the
 * field g of a field that never started stays untouched. Returns 0 when every field converged,
 * 1 when one did not (not converged or stalled), or the first negative error in field order. */
int dmtz_correct_batch_f32(dmtz_ctx* ctx, const dmtz_field_f32* fields, int32_t nfields, const dmtz_flags* flags,
                           dmtz_counts* counts_out, int32_t* status_out, int32_t* sweeps_out);
int dmtz_correct_batch_f64(dmtz_ctx* ctx, const dmtz_field_f64* fields, int32_t nfields, const dmtz_flags* flags,
                           dmtz_counts* counts_out, int32_t* status_out, int32_t* sweeps_out);

/* The classifier alone: critical cells of one field (host pointer) under the reference's
 * five-pass gradient (main_3d.cu:1930-2316), i.e. the cells left unpaired (-1) in paired_vertices
 * / paired_edges / paired_faces / paired_tetrahedras. counts4 = {minima, 1-saddles, 2-saddles,
 * maxima}. crit_lo / crit_hi (nullable, N entries each) receive per vertex the critical cells it
 * owns as a 75-bit mask over its local cells: bit 0 = the vertex, bits 1..14 = the edge to
 * neighbour slot a (1 + a), bits 15..50 = star triangle t (15 + t) in crit_lo, and star tet q in
 * bit q of crit_hi; the slots follow the stencil order of main_3d.cu:44-51. */
int dmtz_classify_f32(dmtz_ctx* ctx, const float* field, int64_t nx, int64_t ny, int64_t nz, int64_t* counts4,
                      uint64_t* crit_lo, uint32_t* crit_hi);
int dmtz_classify_f64(dmtz_ctx* ctx, const double* field, int64_t nx, int64_t ny, int64_t nz, int64_t* counts4,
                      uint64_t* crit_lo, uint32_t* crit_hi);

/* No counterpart in the reference: dmtz_classify_f32 for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * d_field, d_crit_lo and d_crit_hi (both nullable) are device pointers on ctx's device; counts4
 * is host memory; the outputs mean what they mean for dmtz_classify_f32. The call allocates
 * nothing: its scratch (star codes and counters) is a fixed device array of the library, 64 MiB
 * per GPU for at most 2^21 vertices (128^3), held until the process ends; larger grids return
 * DMTZ_E_CAPACITY. All calls on one GPU share that array, so calls that overlap on one GPU
 * (from different contexts or threads) race. */
int dmtz_classify_noalloc_f32(dmtz_ctx* ctx, const float* d_field, int64_t nx, int64_t ny, int64_t nz,
                              int64_t* counts4, uint64_t* d_crit_lo, uint32_t* d_crit_hi);

/* Edit stream of a correction: the reference's encodings of what a correction changed, built on
 * the device from the decompressed field d_g0, the corrected field d_g1 and (edit_type 1) the
 * counter output d_counter, n values each (device pointers on ctx's device).
 *   edit_type 1, cost (main_3d.cu:1494-1646): d_index receives a 3-bit index per vertex packed
 *     least significant bit first (encodeTo3BitBitmap, main_3d.cu:1446-1476; (3n + 7) / 8 bytes):
 *     0 = unchanged, k = 1..5 = the k - 1 steps -2^-s eps of the counter-driven edit, 6 = value
 *     carried in d_edits; d_edits receives g1 - g0 (double) per index 6 in vertex order (the
 *     reference stores (f - eps) - g0 there, the same double for its snap, main_3d.cu:1507).
 *   edit_type 0, original_cost (main_3d.cu:1649-1757): d_index receives the changed vertices'
 *     indices as int32 differences (the first index, then the gap to the previous), 4 bytes each;
 *     d_edits receives g1 - g0 per changed vertex.
 * A changed vertex whose counter does not replay to its value (counters reset by filtration,
 * main_3d.cu:6782; learned warm edits; the finisher's counter 7) is written as index 6 with its
 * value, so every stream decodes back to d_g1 with dmtz_edit_apply_*. out receives the sizes:
 * edited = changed vertices (the reference's edit_ratio is edited / n, main_3d.cu:1553),
 * index_bytes and n_edits. When index_cap or edits_cap is below them (or a buffer is null) the
 * sizes are still returned with DMTZ_E_CAPACITY. */
typedef struct dmtz_edit_stream {
    int64_t edited, index_bytes, n_edits;
} dmtz_edit_stream;
int dmtz_edit_stream_f32(dmtz_ctx* ctx, const float* d_g0, const float* d_g1, const uint8_t* d_counter, int64_t n,
                         double eps, int32_t edit_type, uint8_t* d_index, int64_t index_cap, double* d_edits,
                         int64_t edits_cap, dmtz_edit_stream* out);
int dmtz_edit_stream_f64(dmtz_ctx* ctx, const double* d_g0, const double* d_g1, const uint8_t* d_counter, int64_t n,
                         double eps, int32_t edit_type, uint8_t* d_index, int64_t index_cap, double* d_edits,
                         int64_t edits_cap, dmtz_edit_stream* out);
/* Decoding of an edit stream: d_g holds the decompressed field on entry and the corrected field on
 * return (no counterpart in the reference, which only writes the streams). */
int dmtz_edit_apply_f32(dmtz_ctx* ctx, float* d_g, int64_t n, double eps, int32_t edit_type, const uint8_t* d_index,
                        int64_t index_bytes, const double* d_edits, int64_t n_edits);
int dmtz_edit_apply_f64(dmtz_ctx* ctx, double* d_g, int64_t n, double eps, int32_t edit_type, const uint8_t* d_index,
                        int64_t index_bytes, const double* d_edits, int64_t n_edits);

/* Run-length code of a bitmap on the device (runLengthEncodeBitmap main_3d.cu:1399-1426): the
 * bits of every byte from the most significant down, as (bit, run length) pairs of int32 in
 * d_pairs (2 * cap_runs ints); an empty bitmap gives no runs. n_runs receives the run count;
 * DMTZ_E_CAPACITY when cap_runs is smaller (call with d_pairs = NULL to size it). At most 2^28 - 1
 * bytes, so that every run length fits the reference's int. */
int dmtz_bitmap_rle(dmtz_ctx* ctx, const uint8_t* d_bitmap, int64_t nbytes, int32_t* d_pairs, int64_t cap_runs,
                    int64_t* n_runs);

/* Field files. getdata2 (main_3d.cu:153-181): a raw float64 file without header, in host byte
 * order (little endian here); n_out receives the number of doubles, and out (nullable, cap
 * doubles) the values. saveVectorToBin (main_3d.cu:1315-1327): an 8-byte count (size_t) followed
 * by the doubles; dmtz_read_bin_f64 reads such a file back. DMTZ_E_IO when a file cannot be
 * opened, read or written or is malformed; DMTZ_E_CAPACITY when cap is too small (n_out is still
 * set). */
int dmtz_read_raw_f64(const char* path, double* out, int64_t cap, int64_t* n_out);
int dmtz_write_bin_f64(const char* path, const double* data, int64_t n);
int dmtz_read_bin_f64(const char* path, double* out, int64_t cap, int64_t* n_out);

/* Per-sweep counts of the last correction on ctx: 8 int64 per sweep (dmtz_counts order), up to
 * cap sweeps; returns the number of sweeps recorded. Mirrors the per-sweep prints of
 * main_3d.cu:6787-6791, 7146-7147. */
int64_t dmtz_trajectory(const dmtz_ctx* ctx, int64_t* out, int64_t cap);

/* Phase times of the last correction in seconds, the reference's phase names
 * (main_3d.cu:6419-6435): out[0] gradient_pairing (classification), [1] get_f_cp,
 * [2] get_f_path, [3] fix_f_cp, [4] fix_f_vpath, [5] total. */
int dmtz_phase_times(const dmtz_ctx* ctx, double* out6);

/* Live statistics of the dense mismatch-detection kernel (the per-sweep classification of
 * main_3d.cu:6864-6874 + 6903-6909) accumulated over calls on ctx, measured with HIP events on
 * the context stream: out4 = {launches, total milliseconds, bytes the kernel must move (a 4-byte
 * count per 2048-voxel tile, the 16-byte mismatch masks of the tiles whose count is nonzero,
 * and 8 bytes written per emitted mismatch record), voxels classified}. reset != 0 zeroes
 * them. */
int dmtz_kernel_stats(dmtz_ctx* ctx, double* out4, int reset);

/* Outcome of the last correction on ctx beyond the counts: out4 = {stalls, vertices changed by
 * the finisher, finisher radius reached, lower stars re-evaluated incrementally}. No reference
 * counterpart (the reference neither detects stalls nor finishes them). */
int dmtz_run_info(const dmtz_ctx* ctx, int64_t* out4);

/* The reference's pairing arrays of one field (main_3d.cu:1930-2316, the five passes of
 * main_3d.cu:6550-6564): host arrays of N, E, F, T int64 entries (reference simplex IDs,
 * -1 = critical). Evaluated on the device from per-vertex lower stars. */
int dmtz_gradient_f32(dmtz_ctx* ctx, const float* field, int64_t nx, int64_t ny, int64_t nz,
                      int64_t* pv, int64_t* pe, int64_t* pf, int64_t* pt);
int dmtz_gradient_f64(dmtz_ctx* ctx, const double* field, int64_t nx, int64_t ny, int64_t nz,
                      int64_t* pv, int64_t* pe, int64_t* pf, int64_t* pt);

/* One detection + repair phase on (f, g) without applying it: the mismatch counts of
 * main_3d.cu:2492-2550 (counts4 = min, 1-saddle, 2-saddle, max) and the edit set the repair
 * kernels of main_3d.cu:2554-4303 write into d_deltaBuffer (targets: N bytes, 0/1). */
int dmtz_targets_f32(dmtz_ctx* ctx, const float* f, const float* g, int64_t nx, int64_t ny, int64_t nz,
                     double eps, double range, const dmtz_flags* flags, uint8_t* targets, int64_t* counts4);
int dmtz_targets_f64(dmtz_ctx* ctx, const double* f, const double* g, int64_t nx, int64_t ny, int64_t nz,
                     double eps, double range, const dmtz_flags* flags, uint8_t* targets, int64_t* counts4);

/* Separatrix detection on (f, g) as they stand (main_3d.cu:4437-5062 in the mode the flags
 * select: geometry = get_false_as_vpath / get_false_ds_vpath, connectivity = the _connectivity
 * variants with f's endpoint labels of main_3d.cu:4613-4759). counts2 = {descending, ascending};
 * ds_id/ds_rec receive (f-critical edge, recorded vertex) and as_id/as_rec (tet, divergent tet)
 * pairs, as many as the caller's capacity cap allows (same order). When targets is non-null it
 * receives the edit set of fix_as_vpath / fix_ds_vpath (main_3d.cu:5135-5295) over those lists. */
int dmtz_paths_f32(dmtz_ctx* ctx, const float* f, const float* g, int64_t nx, int64_t ny, int64_t nz,
                   const dmtz_flags* flags, int64_t* counts2, int64_t* ds_id, int64_t* ds_rec, int64_t* as_id,
                   int64_t* as_rec, int64_t cap, uint8_t* targets);
int dmtz_paths_f64(dmtz_ctx* ctx, const double* f, const double* g, int64_t nx, int64_t ny, int64_t nz,
                   const dmtz_flags* flags, int64_t* counts2, int64_t* ds_id, int64_t* ds_rec, int64_t* as_id,
                   int64_t* as_rec, int64_t cap, uint8_t* targets);

/* 2-saddle walls on (f, g) as they stand: get_false_ds_wall (main_3d.cu:5064-5133) made
 * deterministic. From every f-critical face, a depth-first walk over f's pairs (face -> its edges
 * (v1,v2), (v1,v3), (v2,v3) of the descending vertex triple -> the face paired with each edge in
 * f when it pairs back), with its own stack and visited set, records the first edge whose f and
 * g partners differ. count receives the number of such faces; face/edge receive (face id, edge
 * id) pairs, as many as cap allows (in no particular order). When targets is non-null it
 * receives the edit set of fix_ds_wall (main_3d.cu:5297-5413) over the list; eps and range enter
 * its threshold, and flags (nullable) its edit_type. DMTZ_E_CAPACITY if one wall walk visits more
 * than 2^17 faces. */
int dmtz_walls_f32(dmtz_ctx* ctx, const float* f, const float* g, int64_t nx, int64_t ny, int64_t nz, double eps,
                   double range, const dmtz_flags* flags, int64_t* count, int64_t* face, int64_t* edge, int64_t cap,
                   uint8_t* targets);
int dmtz_walls_f64(dmtz_ctx* ctx, const double* f, const double* g, int64_t nx, int64_t ny, int64_t nz, double eps,
                   double range, const dmtz_flags* flags, int64_t* count, int64_t* face, int64_t* edge, int64_t cap,
                   uint8_t* targets);

#ifdef __cplusplus
}
#endif

#endif /* DMTZ_H */
