/*
 * C ABI of the learned SCF starting density (part of libvlxfock.so).
 *
 * No reference counterpart: the reference starts the SCF from the
 * superposition-of-atomic-densities guess (src/pymodule/sadguessdriver.py:841-971,
 * called from src/pymodule/scfdriver.py:777-815), which stays the fallback.
 * A small network predicts each atom-pair block of the density from
 * rotation-invariant pair features; it is trained on converged densities that
 * this library's own SCF produces.
 *
 * The network: n_in -> n_hidden -> n_hidden -> n_out, SiLU after the two hidden
 * layers, BF16 weights and activations on the matrix cores
 * (v_mfma_f32_16x16x32_bf16), FP32 accumulation, FP32 master weights and Adam.
 * Parameters are packed as W0 (n_in x n_hidden, row-major), b0, W1
 * (n_hidden x n_hidden), b1, W2 (n_hidden x n_out), b2.  All batch arrays are
 * device buffers (HBM of the context's device), row-major, one sample per row;
 * work runs on the context's stream.  Return codes as in vlxfock.h.
 */
#ifndef VLXGUESS_H
#define VLXGUESS_H

#include "vlxfock.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct vlx_mlp vlx_mlp;

/* A network with batches of up to max_batch samples; an output row is a
 * density block flattened with block_cols columns (the loss masks rows >= ni
 * and columns >= nj of sample (ni, nj)).  Parameters start at zero: set them
 * with vlx_mlp_set_params. */
int vlx_mlp_create(vlx_context* ctx, int n_in, int n_hidden, int n_out, int block_cols, int max_batch,
                   vlx_mlp** out);
void vlx_mlp_destroy(vlx_mlp* net);

/* Number of packed parameters. */
int vlx_mlp_num_params(vlx_mlp* net, long long* n);

/* Host copies of the packed FP32 parameters in and out (set also restarts
 * the optimiser; both synchronise the stream). */
int vlx_mlp_set_params(vlx_mlp* net, const float* params);
int vlx_mlp_get_params(vlx_mlp* net, float* params);

/* d_y (m x n_out, FP32) = network(d_x (m x n_in, FP32, rounded to BF16 on
 * input)).  No synchronisation. */
int vlx_mlp_forward(vlx_mlp* net, const float* d_x, int m, float* d_y);

/* Forward pass, loss and gradients of one batch: d_target (m x n_out, FP32),
 * d_shape (m x 2 ints: block rows ni and columns nj of each sample);
 * loss = (1/m) sum_s 1/(ni nj) sum_{r < ni, c < nj} (y - t)^2.  The outputs
 * go to d_y (m x n_out), the loss to *loss (host), the packed FP32 gradient
 * stays on the device for vlx_mlp_adam_step and is copied to grads (host,
 * num_params floats) when grads is not NULL.  Synchronises the stream. */
int vlx_mlp_loss_grad(vlx_mlp* net, const float* d_x, const float* d_target, const int* d_shape, int m, float* d_y,
                      float* loss, float* grads);

/* One Adam step with the last gradient (bias-corrected moments). */
int vlx_mlp_adam_step(vlx_mlp* net, float lr, float beta1, float beta2, float eps);

#ifdef __cplusplus
}
#endif

#endif /* VLXGUESS_H */
