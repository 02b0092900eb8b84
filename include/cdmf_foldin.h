/*
 * C ABI of the fold-in extension: folding new rows (users) into a trained
 * factorization with the item factors H fixed, from zero (plain CCD++
 * fold-in) or from a learned warm start (a small MLP that maps features of
 * the row to an initial w_i).  No counterpart in the reference, which only
 * factorizes whole matrices (src/cdmf_ref.cpp:47-140); the per-rank update is
 * the u-side update of src/cdmf_ref.cpp:7-17 applied to one row.
 *
 * Conventions as in cdmf_hip.h: every call returns 0 or a negative
 * CCD_ERR_* code; host arrays are borrowed for the call; objects own their
 * device memory and one HIP stream each.
 */
#ifndef CDMF_FOLDIN_H
#define CDMF_FOLDIN_H

#include <stdint.h>

#include "cdmf_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- MLP --- */
/* in -> hidden -> hidden -> out, GELU after both hidden layers.  Parameters,
 * flat FP32: W1 [hidden][in], b1 [hidden], W2 [hidden][hidden], b2 [hidden],
 * W3 [out][hidden], b3 [out].  Inputs are standardised with in_mean/in_std
 * and outputs mapped back with out_mean/out_std (ccd_mlp_set_norm; identity
 * until set).  Trained in FP32 with Adam.  Inference (bf16_weights below):
 * 0 = FP32; 1 = weights read as bfloat16 (round to nearest even),
 * arithmetic in FP32; 2 = the matrix-core path: one launch on the bfloat16
 * MFMA units, inputs, weights and hidden activations rounded to bfloat16,
 * products accumulated in FP32 (needs 256 hidden units, in <= 288,
 * out <= 256). */
typedef struct ccd_mlp ccd_mlp;

int32_t ccd_mlp_create(int32_t device, uint32_t in, uint32_t hidden, uint32_t out, uint64_t seed, ccd_mlp** m);
int32_t ccd_mlp_destroy(ccd_mlp* m);
const char* ccd_mlp_last_error(ccd_mlp* m);
uint64_t ccd_mlp_param_count(ccd_mlp* m);
int32_t ccd_mlp_get_params(ccd_mlp* m, float* params);
int32_t ccd_mlp_set_params(ccd_mlp* m, const float* params);
int32_t ccd_mlp_set_norm(ccd_mlp* m, const float* in_mean, const float* in_std, const float* out_mean,
                         const float* out_std);
/* Y [rows][out] = MLP(X [rows][in]) (raw features in, raw outputs out);
 * bf16_weights 0, 1 or 2 as above. */
int32_t ccd_mlp_forward(ccd_mlp* m, uint64_t rows, const float* X, float* Y, int32_t bf16_weights);
/* Loss mean((MLP_n(Xn) - Tn)^2) and its gradient (flat, like the
 * parameters) on already standardised inputs Xn and targets Tn, FP32
 * weights; no update.  For gradient checks. */
int32_t ccd_mlp_loss_grad(ccd_mlp* m, uint64_t rows, const float* Xn, const float* Tn, double* loss, float* grad);
/* Adam on (X [rows][in], T [rows][out]) raw pairs, standardised with the
 * norms: `epochs` passes in seeded random order, minibatches of `batch`
 * rows.  *epoch_loss = mean minibatch loss of the last epoch. */
int32_t ccd_mlp_train(ccd_mlp* m, uint64_t rows, const float* X, const float* T, uint32_t epochs, uint32_t batch,
                      float lr, uint64_t seed, double* epoch_loss);

/* ------------------------------------------------------------ fold-in --- */
typedef struct ccd_fold ccd_fold;

/* H [k][n] rank-major (as ccd_get_factors returns it), k <= 128. */
int32_t ccd_fold_create(int32_t device, uint32_t k, int64_t n, const float* H, float lambda, ccd_fold** f);
int32_t ccd_fold_destroy(ccd_fold* f);
const char* ccd_fold_last_error(ccd_fold* f);
/* F [rows][2k + 2] of CSR rows (ptr[0] = 0): the rating-weighted mean and
 * variance of the rated items' H columns, log(1 + degree), mean rating. */
int32_t ccd_fold_features(ccd_fold* f, int64_t rows, const uint32_t* ptr, const uint32_t* idx, const float* val,
                          float* F);
/* W [rows][k] after `sweeps` CCD++ sweeps over every row with H fixed,
 * starting from W0 ([rows][k]; NULL = zeros).  *seconds: device time. */
int32_t ccd_fold_rows(ccd_fold* f, int64_t rows, const uint32_t* ptr, const uint32_t* idx, const float* val,
                      const float* W0, int32_t sweeps, float* W, double* seconds);
/* ccd_fold_rows without any device allocation: no counterpart in the
 * reference; for callers that cannot allocate (serving loops folding in
 * batches of new users, code that must not call hipMalloc).  Same kernels
 * and result as ccd_fold_rows (bit for bit); the batch is staged in a
 * library-owned static device scratch of about 28 MB per GPU, held until
 * the process ends, so one call takes at most 4096 rows and 2,097,152
 * ratings (larger batches fail with CCD_ERR_INVALID and a message).  Every
 * call on one GPU shares that GPU's scratch: calls that overlap on one GPU
 * (from several threads or fold objects) race. */
int32_t ccd_fold_rows_noalloc(ccd_fold* f, int64_t rows, const uint32_t* ptr, const uint32_t* idx, const float* val,
                              const float* W0, int32_t sweeps, float* W, double* seconds);
/* Warm start: features -> MLP (bf16_weights 0, 1 or 2 as for
 * ccd_mlp_forward) -> W0,
 * then `sweeps` sweeps; the MLP must map 2k + 2 inputs to k outputs.
 * *seconds: device time of the whole pipeline. */
int32_t ccd_fold_warm(ccd_fold* f, ccd_mlp* m, int64_t rows, const uint32_t* ptr, const uint32_t* idx,
                      const float* val, int32_t sweeps, int32_t bf16_weights, float* W, double* seconds);

#ifdef __cplusplus
}
#endif

#endif /* CDMF_FOLDIN_H */
