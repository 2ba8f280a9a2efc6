/* axicor_surrogate.h — C-ABI of libaxicor_surrogate.so: the MLP surrogate that
 * proposes GLIF/GSOP parameter sets for the parameter search.
 *
 * No function here replaces a reference symbol. The reference searches with
 * Optuna TPE over a live brain (examples/bayesian_search.py:63-187). The
 * surrogate is trained on (parameter, outcome) pairs produced by the batched
 * Day-Phase engine (axc_batch_* in axicor_day.h). It scores candidate
 * parameter vectors so that the exact simulator only judges the most promising
 * ones.
 *
 * Network: d_in -> hidden x depth -> d_out, GELU (tanh form) after each hidden
 * layer, linear output. Weights have fp32 master copies and bf16 copies. Inputs
 * and activations are bf16; GEMMs accumulate in fp32 on gfx950 MFMA. Training
 * minimises the mean squared error with Adam (b1 0.9, b2 0.999, eps 1e-8,
 * bias-corrected).
 *
 * Sizes are padded inside: d_in to a multiple of 32, hidden to a multiple of
 * 32, d_out to a multiple of 16. Padding entries are zero and stay zero.
 * Parameter arrays exchanged with the host (axc_mlp_get_params, set_params,
 * gradients) use the padded layout, layer by layer: W[in_pad][out_pad]
 * row-major, then b[out_pad]. axc_mlp_layout reports the padded sizes.
 *
 * Pointers named d_* are device pointers; the others are host pointers.
 * Streams are hipStream_t passed as void* (0 = the null stream). Functions
 * return 0 or a hipError_t value; invalid arguments give hipErrorInvalidValue
 * (1). A handle must not be used by two threads at once.
 */
#ifndef AXICOR_SURROGATE_H
#define AXICOR_SURROGATE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Creates a network with He-uniform initial weights drawn from `seed`, zero
 * biases and buffers for batches of up to max_batch rows. */
int32_t axc_mlp_create(int d_in, int hidden, int depth, int d_out, int max_batch, uint64_t seed, void** out);

/* d_y[n][d_out] = net(d_x[n][d_in]) in fp32. Batches above max_batch are
 * processed in chunks. Asynchronous on `stream`. */
int32_t axc_mlp_forward(void* mlp, const float* d_x, int64_t n, float* d_y, void* stream);

/* One Adam step on the mean squared error of n <= max_batch rows against
 * d_t[n][d_out]. If `loss` is given, the batch loss before the step is
 * written to it, which synchronises the stream. */
int32_t axc_mlp_train_step(void* mlp, const float* d_x, const float* d_t, int n, float lr, float* loss, void* stream);

/* Gradients of the batch loss without a step, copied to host_grads in the
 * padded parameter layout; the loss goes to `loss` if given. Synchronous. */
int32_t axc_mlp_gradients(void* mlp, const float* d_x, const float* d_t, int n, float* host_grads, float* loss,
                          void* stream);

/* Padded layer sizes (arrays of at least depth+1 entries) and the total
 * parameter count of the padded layout. Any output may be NULL. */
int32_t axc_mlp_layout(void* mlp, int32_t* n_layers, int32_t* ins, int32_t* outs, int64_t* n_params);

/* fp32 master parameters to / from a host array in the padded layout.
 * set_params also refreshes the bf16 copies. Padding entries given to
 * set_params must be zero. Both synchronise the device. */
int32_t axc_mlp_get_params(void* mlp, float* host);
int32_t axc_mlp_set_params(void* mlp, const float* host);

/* Clears the Adam moments and the step count. */
int32_t axc_mlp_reset_optimizer(void* mlp);

/* The network's GEMM kernel on its own: d_c[M][ldc] (fp32) = op(A) op(B),
 * with bf16 inputs. op(A)(m,k) = A[m*lda+k], or A[k*lda+m] if ta.
 * op(B)(k,n) = B[k*ldb+n], or B[n*ldb+k] if tb. lda and ldb must be
 * multiples of 8. Asynchronous. */
int32_t axc_mlp_gemm(int M, int N, int K, int ta, int tb, const void* d_a, int lda, const void* d_b, int ldb,
                     float* d_c, int ldc, void* stream);

/* Whether axc_mlp_forward runs the whole network in one kernel (128 rows
 * per workgroup, activations kept in LDS; on by default) or layer by layer.
 * Both give the same outputs bit for bit. The fused kernel serves networks
 * with d_in <= 32, hidden 481..512, d_out <= 16 and at most 7 hidden layers;
 * others always run layer by layer. Returns 1 if the network has that
 * shape, 0 if not (hipErrorInvalidValue for a null handle). */
int32_t axc_mlp_set_fused_forward(void* mlp, int enable);

/* Waits for the device and frees every buffer of the network. */
void axc_mlp_destroy(void* mlp);

#ifdef __cplusplus
}
#endif

#endif /* AXICOR_SURROGATE_H */
