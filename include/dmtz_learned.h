/* dmtz_learned.h - the learned warm start of the topology correction.
 *
 * No counterpart in the reference (main_3d.cu has only the plain loop). A small 3-D convolutional
 * network predicts, from the decompressed field g alone, which vertices the plain loop will edit
 * and by how much; the prediction is applied as a first edit of g (never past the snap of
 * f - eps, so |g - f| <= eps still holds), and the plain loop (dmtz_correct_*) then runs from
 * there and decides the result: it has zero topology mismatches exactly as without the
 * network.
 *
 * Network: four 3x3x3 convolutions with zero padding, 16 channels (8 used input channels,
 * 16-16-16 hidden with ReLU, 2 outputs: edit logit and edit size in units of eps), bf16 weights
 * and activations, fp32 accumulation on the matrix cores (v_mfma_f32_16x16x16_bf16), fp32 master
 * parameters trained with Adam by the library's own forward and backward kernels.
 *
 * Per-vertex input channels (bf16, 16 per vertex, x fastest, then y, then z):
 *   0: (g(v) - mean of g over the 14 stencil neighbours) / eps, clamped to [-8, 8]
 *   1..4: critical cells of g owned by v (vertex, edges, triangles, tets), as in dmtz_classify
 *   5: log10(eps / range) + 3
 *   6: (g(v) - min of the neighbours) / eps, clamped to [-8, 8]
 *   7: (max of the neighbours - g(v)) / eps, clamped to [-8, 8]
 *   8..15: zero
 * Labels (fp32, 2 per vertex): y = 1 if the plain loop changed v, t = (g - g') / eps.
 * Loss: pos_weight-weighted binary cross entropy of the logit against y, plus
 * lambda * y * (size - t)^2, both averaged over the vertices.
 *
 * Parameters: 4 layers of W[27][16][16] (tap = (dz+1)*9 + (dy+1)*3 + (dx+1), input channel,
 * output channel) followed by b[16]; 6928 floats per layer, DMTZ_NET_PARAMS in all. Entries of
 * unused channels (layer 0 inputs 8..15, layer 3 outputs 2..15) stay zero. */
#ifndef DMTZ_LEARNED_H
#define DMTZ_LEARNED_H

#include <stdint.h>

#include "dmtz.h"

#ifdef __cplusplus
extern "C" {
#endif

#define DMTZ_NET_CHANNELS 16
#define DMTZ_NET_LAYER_PARAMS 6928
#define DMTZ_NET_PARAMS (4 * DMTZ_NET_LAYER_PARAMS)

typedef struct dmtz_net dmtz_net;

/* A network on the context's device, He-initialised from seed. */
int dmtz_net_create(dmtz_ctx* ctx, uint64_t seed, dmtz_net** out);
void dmtz_net_destroy(dmtz_net* net);
/* Host copies of the fp32 master parameters (DMTZ_NET_PARAMS floats). set_params also resets
 * the optimizer state. */
int dmtz_net_params(dmtz_ctx* ctx, dmtz_net* net, float* out);
int dmtz_net_set_params(dmtz_ctx* ctx, dmtz_net* net, const float* in);

/* Input channels of device field d_g into d_feat (device, 16 bf16 per vertex). */
int dmtz_net_features_f32(dmtz_ctx* ctx, const float* d_g, int64_t nx, int64_t ny, int64_t nz, double eps,
                          double range, uint16_t* d_feat);
/* Forward pass: d_out (device) receives 2 floats per vertex (edit logit, edit size / eps). */
int dmtz_net_forward(dmtz_ctx* ctx, dmtz_net* net, const uint16_t* d_feat, int64_t nx, int64_t ny, int64_t nz,
                     float* d_out);
/* Forward and backward pass without an update: grads_out (host, nullable) receives the gradient
 * of the loss with respect to the DMTZ_NET_PARAMS parameters, loss_out (host, nullable) the two
 * loss terms. */
int dmtz_net_gradient(dmtz_ctx* ctx, dmtz_net* net, const uint16_t* d_feat, const float* d_labels, int64_t nx,
                      int64_t ny, int64_t nz, float pos_weight, float lambda, float* grads_out, double* loss_out);
/* One Adam step (beta1 0.9, beta2 0.999) of learning rate lr on one field. */
int dmtz_net_train_step(dmtz_ctx* ctx, dmtz_net* net, const uint16_t* d_feat, const float* d_labels, int64_t nx,
                        int64_t ny, int64_t nz, float pos_weight, float lambda, float lr, double* loss_out);

/* Warm-started correction of device fields: features of g, forward pass, then every vertex with
 * a positive edit logit gets g = max(g - size * eps, snap(f - eps)) with size clamped to [0, 2]
 * (warm_edits_out, nullable, receives their number); then the plain loop of
 * dmtz_correct_dev_f32 with the given flags, which returns as there. In d_counter_out the
 * vertices the warm start moved carry counter 6, the edit-stream value whose edit is stored
 * explicitly (main_3d.cu:1502-1508), unless the loop left them at 6 or 7 already. */
int dmtz_correct_learned_f32(dmtz_ctx* ctx, dmtz_net* net, const float* d_f, float* d_g, int64_t nx, int64_t ny,
                             int64_t nz, double eps, double range, const dmtz_flags* flags, dmtz_counts* out,
                             uint8_t* d_counter_out, int32_t* sweeps_out, int64_t* warm_edits_out);

/* The batched form with the learned warm start: each field of the array (device pointers only,
 * on_device != 0) gets the warm edit of dmtz_correct_learned_f32, one field after another on the
 * context's stream, and then all fields run through dmtz_correct_batch_f32 with the given flags,
 * which returns as there. warm_edits_out (nullable) receives one count per field. */
int dmtz_correct_batch_learned_f32(dmtz_ctx* ctx, dmtz_net* net, const dmtz_field_f32* fields, int32_t nfields,
                                   const dmtz_flags* flags, dmtz_counts* counts_out, int32_t* status_out,
                                   int32_t* sweeps_out, int64_t* warm_edits_out);

#ifdef __cplusplus
}
#endif

#endif /* DMTZ_LEARNED_H */
