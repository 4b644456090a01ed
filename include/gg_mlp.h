/*
 * gg_mlp.h -- the swap-ranking surrogate network of libggeval.so: a small
 * MLP (ReLU hidden layers, linear output) trained and evaluated on the GPU
 * with BF16 MFMA GEMMs (FP32 accumulate) and Adam on FP32 master weights.
 *
 * No reference counterpart: the reference draws 2-opt proposals uniformly
 * (general/sa.c:217-369).  The network only orders proposals; the exact
 * evaluator still scores every move that is accepted.
 *
 * Parameter layout (flat FP32 array): for layer i, W_i [rows_i][cols_i]
 * row-major (cols_i = output width rounded up to 32; rows_0 = input width
 * rounded up to 8, rows_i = cols_{i-1}; the extra columns and rows being
 * zero) followed by b_i [cols_i]; see gg_mlp_layout().
 *
 * GEMMs: gfx950 v_mfma_f32_32x32x16_bf16, 64x64 output tiles per 256-thread
 * workgroup, operands staged through LDS as 16-byte vectors.
 */
#ifndef GG_MLP_H
#define GG_MLP_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct gg_mlp gg_mlp;

/* dims[0..n_layers]: input width, hidden widths, output width. */
gg_mlp *gg_mlp_create(int n_layers, const int *dims, uint64_t seed);
void gg_mlp_destroy(gg_mlp *m);
int gg_mlp_param_count(const gg_mlp *m);
/* Fills per-layer rows, padded columns, weight and bias offsets; returns the layer count. */
int gg_mlp_layout(const gg_mlp *m, int *rows, int *cols_padded, int *w_off, int *b_off);
int gg_mlp_get_params(gg_mlp *m, float *out);
int gg_mlp_set_params(gg_mlp *m, const float *in);
/* y[i] = network(x[i]) for m rows of x (host FP32 [m][input width]). */
int gg_mlp_forward(gg_mlp *m, const float *x, int rows, float *y);
/* Mean squared error against t and its gradient w.r.t. every parameter
 * (grad may be NULL); no update. */
int gg_mlp_gradient(gg_mlp *m, const float *x, const float *t, int rows, float *loss, float *grad);
/* One Adam step on the MSE loss; *loss receives the pre-update loss. */
int gg_mlp_train_step(gg_mlp *m, const float *x, const float *t, int rows, float lr, float *loss);

#ifdef __cplusplus
}
#endif
#endif
