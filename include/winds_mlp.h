/*
 * winds_mlp.h -- far-wake surrogate network: a multilayer perceptron with
 * SiLU hidden layers, BF16 weights and activations, FP32 accumulation on the
 * MFMA units, FP32 master weights and Adam for training.
 *
 * No counterpart exists in the reference (WInDS evaluates the whole wake by
 * direct summation, core/BiotSavart.m); the network replaces the far wake's
 * induced velocity at near-wake targets (winds_march.h surrogate mode).
 *
 * All data pointers are device memory, row-major: inputs [B][n_in] and
 * targets / outputs [B][n_out], FP32.  Parameters on the host are one flat
 * FP32 array: for each layer W [n_out_l][n_in_l] then b [n_out_l].
 */
#ifndef WINDS_MLP_H
#define WINDS_MLP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct winds_mlp winds_mlp;

/* n_hidden_layers layers of width n_hidden (<= 512) between n_in (<= 256)
 * inputs and n_out (<= 32) outputs; He-initialised from seed. */
int winds_mlp_create(int32_t n_in, int32_t n_hidden, int32_t n_hidden_layers, int32_t n_out,
                     int32_t max_batch, uint64_t seed, winds_mlp **out);
void winds_mlp_destroy(winds_mlp *m);

/* Number of parameters of the flat host layout. */
int64_t winds_mlp_num_params(const winds_mlp *m);

/* out = {n_in, n_hidden, n_hidden_layers, n_out, max_batch (rounded up to a
 * multiple of 64), parameters}. */
int winds_mlp_info(const winds_mlp *m, int64_t out[6]);
int winds_mlp_get_params(winds_mlp *m, float *host);
int winds_mlp_set_params(winds_mlp *m, const float *host);

/* Y = net(X) for B rows (any B <= max_batch), enqueued on stream (NULL: the
 * default stream).  Calls on one network share its activation buffers, so
 * they must not run concurrently on different streams. */
int winds_mlp_forward(winds_mlp *m, const float *X, int32_t B, float *Y, void *stream);

/* Forward path of winds_mlp_forward and of the training step (no
 * counterpart in the reference, which has no network): on = 1 runs the
 * whole forward pass in one kernel per 32 rows, activations held in LDS
 * (the default where the hidden width is at most 256); on = 0 runs one GEMM
 * per layer.  Both give bitwise the same outputs, gradients and updates.
 * WINDS_E_DIMS for on = 1 with a hidden width above 256. */
int winds_mlp_set_fused(winds_mlp *m, int32_t on);

/* Mean-squared-error gradients of all parameters for one batch (no update),
 * copied to host in the flat layout; *loss receives the batch loss.  Runs on
 * the network's own stream: X and T must be ready when it is called. */
int winds_mlp_gradients(winds_mlp *m, const float *X, const float *T, int32_t B, float *grads,
                        float *loss);

/* One Adam step on the MSE loss of the batch (B <= max_batch) on stream
 * (NULL: the default stream).  loss (device FP32 scalar, may be NULL)
 * receives the batch loss before the update.  From the second of
 * consecutive calls with the same X, T, B and stream, the step (about 30
 * kernels) is recorded once into a HIP graph, on the network's own stream,
 * and replayed on stream; X and T are read when each step runs. */
int winds_mlp_train_step(winds_mlp *m, const float *X, const float *T, int32_t B, float lr,
                         float *loss, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* WINDS_MLP_H */
