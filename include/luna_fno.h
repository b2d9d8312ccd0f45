/*
 * luna_fno.h -- C ABI of the spectral surrogate (1-D Fourier neural operator)
 * in libluna_hip.so.
 *
 * No counterpart in the reference: Luna.jl has no learned model.  The
 * surrogate screens points of a parameter scan (Scans.runscan,
 * src/Scans.jl:275-430): it predicts the exit spectrum of a propagation from
 * (energy, pressure, input spectrum), and points whose predicted error is above
 * a bound are run on the FP64 solver (lh_batch_solve), which stays the judge.
 *
 * Network (Li et al., "Fourier Neural Operator for Parametric PDEs", ICLR 2021):
 *   v0 = P a + bP                                   (lift, per sample point x)
 *   z_l = K_l(v_l) + W_l v_l + b_l,  v_{l+1} = gelu(z_l) (l < L-1), v_L = z_{L-1}
 *   K_l(v) = irfft_X( R_l[m] · rfft_X(v)[m], m < modes )   (complex R_l[m]: C×C)
 *   y = Q2 gelu(Q1 v_L + bq1) + bq2
 * The lift and the Fourier layers (their truncated DFTs, per-mode complex
 * channel mixing and pointwise channel mixing) run on BF16 matrix cores
 * (v_mfma_f32_16x16x32_bf16) with FP32 accumulation, weights held as BF16 for
 * the products and as an FP32 master copy for the AdamW update.  The
 * projection head (Q1, Q2), whose output is the log spectrum itself, runs on
 * FP32 matrix cores (v_mfma_f32_16x16x4f32).  Activations are FP32.
 *
 * Parameter vector layout (FP32, this order):
 *   P[C][cin], bP[C],
 *   per layer: W[C][C] (out, in), b[C], Rre[modes][C][C], Rim[modes][C][C] (mode, in, out),
 *   Q1[H][C], bq1[H], Q2[H], bq2[1].
 * Data layout: inputs x[batch][n_x][cin], outputs/targets y[batch][n_x].
 */
#ifndef LUNA_FNO_H
#define LUNA_FNO_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct lh_fno lh_fno;

typedef struct lh_fno_cfg {
    int32_t n_x;          /* samples per input function (multiple of 64)    */
    int32_t cin;          /* input features per sample point                */
    int32_t channels;     /* C (multiple of 32)                             */
    int32_t modes;        /* retained Fourier modes (<= n_x/2)              */
    int32_t layers;       /* Fourier layers L                               */
    int32_t hidden;       /* H, projection width (multiple of 32)           */
    int32_t max_batch;    /* largest batch for forward / training           */
} lh_fno_cfg;

typedef struct lh_adam {
    float lr, beta1, beta2, eps, weight_decay;  /* AdamW */
} lh_adam;

int lh_fno_create(lh_fno **out, const lh_fno_cfg *cfg);
int lh_fno_destroy(lh_fno *f);
int lh_fno_param_count(const lh_fno *f, int64_t *n);
/* Set the FP32 master parameters (and their BF16 copies); resets Adam state. */
int lh_fno_set_params(lh_fno *f, const float *params);
int lh_fno_get_params(lh_fno *f, float *params);
/* y = net(x) for batch inputs on the host. */
int lh_fno_forward(lh_fno *f, const float *x, int batch, float *y);
/* Weighted mean-squared loss Σ w (y - target)² / (batch · n_x) and its
 * gradient (weight [batch][n_x] may be NULL: unit weights; grad may be NULL). */
int lh_fno_loss_grad(lh_fno *f, const float *x, const float *target, const float *weight, int batch,
                     float *loss, float *grad);
/* Device-resident training set: n samples of x, target and loss weight
 * (weight may be NULL), uploaded once. */
int lh_fno_set_data(lh_fno *f, const float *x, const float *target, const float *weight, int n);
/* n_steps AdamW steps; batch_idx[n_steps*batch] picks samples of the resident
 * set; lr_sched[n_steps] (may be NULL: opt->lr throughout); losses[n_steps]
 * receives each step's loss (may be NULL). */
int lh_fno_train(lh_fno *f, const int32_t *batch_idx, int n_steps, int batch, const lh_adam *opt,
                 const float *lr_sched, float *losses);

#ifdef __cplusplus
}
#endif
#endif /* LUNA_FNO_H */
