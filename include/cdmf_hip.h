/*
 * cdmf_hip.h -- C ABI of the MI355X (gfx950) CCD++ solver.
 *
 * Plain C: pointers, sizes and status codes only.  Host arrays passed in are
 * borrowed for the duration of the call and copied; the context owns all
 * device memory.  A context is bound to one HIP device and one stream and is
 * not thread-safe: drive it from one host thread.
 *
 * Reference interfaces this boundary replaces (all C++ in the reference,
 * src/cdmf.h:4-10, selected by -V in src/main.cpp:38-64):
 *   void cdmf_ocl_01(SparseMatrix&, MatData& W, MatData& H, TestData&, parameter&, char kernel_file[])
 *                                                     src/cdmf_ocl_1.cpp:6-423
 *   void cdmf_ocl_02(SparseMatrix&, MatData& W, MatData& H, TestData&, parameter&, char kernel_file[])
 *                                                     src/cdmf_ocl_2.cpp:6-318
 *   void cdmf_csr5(SparseMatrix&, MatData& W, MatData& H, parameter&, char kernel_file[])
 *                                                     src/cdmf_csr5.cpp:7-369
 * Each of those is one call that uploads R, runs the outer/rank/inner loop
 * and writes W, H back.  Here the same work is split into
 *   ccd_create/ccd_destroy        OpenCL platform/context/queue set-up and release
 *                                 (cdmf_ocl_2.cpp:11-21, 299-317)
 *   ccd_upload                    the clCreateBuffer(COPY_HOST_PTR) uploads of the
 *                                 CSR and CSC copies of R (cdmf_ocl_2.cpp:82-93)
 *   ccd_upload_test               the test-set buffers of V1 (cdmf_ocl_1.cpp:121-140)
 *   ccd_set_factors               the W/H uploads (cdmf_ocl_1.cpp:70-104)
 *   ccd_run                       the oiter/t/inner loop with per-kernel timing
 *                                 and the per-iteration RMSE line (cdmf_ocl_2.cpp:205-297)
 *   ccd_get_factors               the final W/H read-back (cdmf_ocl_1.cpp:388-421)
 *   ccd_get_residual              the residual the oracle leaves in R (cdmf_ref.cpp:120-121)
 *   ccd_rmse                      calculate_rmse_directly on the device (tools.cpp:543-582,
 *                                 GPU_rmse in kcode/ccd_1.cl:1-35)
 * The C++ adapter cdmf_amd/csrc/host/cdmf_hip.cpp restores the one-call
 * signatures above on top of these entry points.
 */
#ifndef CDMF_HIP_H
#define CDMF_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CCD_ABI_VERSION 6

/* Status codes: 0 = success, negative = error (message: ccd_last_error). */
#define CCD_OK 0
#define CCD_ERR_INVALID (-1)     /* bad argument or shape */
#define CCD_ERR_HIP (-2)         /* a HIP runtime call failed */
#define CCD_ERR_ALLOC (-3)       /* device or host allocation failed */
#define CCD_ERR_STATE (-4)       /* call out of order (e.g. run before upload) */
#define CCD_ERR_UNSUPPORTED (-5) /* option not available in this build */

typedef struct ccd_ctx ccd_ctx;

/* Solver variants (-V).  1 and 2 select the degree-binned sweeps (the
 * reference's "Native" and "Thread Batching" kernels have one replacement);
 * 3 selects the tiled segmented-sum sweeps (the CSR5 "Load Balancing" path). */
enum ccd_variant { CCD_VARIANT_NATIVE = 1, CCD_VARIANT_BATCHING = 2, CCD_VARIANT_TILED = 3 };

/* Storage precision of W and H.  Residual values and all arithmetic stay FP32
 * except CCD_PREC_FP64_SUMS, which accumulates each row's numerator and
 * denominator in FP64 (the oracle-check mode). */
enum ccd_precision { CCD_PREC_FP32 = 0, CCD_PREC_BF16 = 1, CCD_PREC_FP64_SUMS = 2 };

/* Which factor is swept first inside a rank.  H_FIRST is the reference
 * (v-sweep over the CSC copy, then the u-sweep, cdmf_ref.cpp:96-107).
 * W_FIRST is the same algorithm on R^T with W and H exchanged. */
enum ccd_order { CCD_ORDER_H_FIRST = 0, CCD_ORDER_W_FIRST = 1 };

typedef struct ccd_params {
    uint32_t k;           /* rank (-k); equals the k given to ccd_set_factors */
    float lambda;         /* regularisation (-l) */
    int32_t maxiter;      /* outer iterations (-t) */
    int32_t maxinneriter; /* inner iterations (-T) */
    int32_t variant;      /* enum ccd_variant */
    int32_t precision;    /* enum ccd_precision */
    int32_t order;        /* enum ccd_order */
    int32_t zero_first;   /* 1: zero the first-swept factor at entry (the reference zeroes H,
                             cdmf_ref.cpp:57-61); 0: keep it (the product W H^T must then be 0) */
    int32_t rmse_every;   /* 1: test RMSE after every outer iteration (needs ccd_upload_test) */
    int32_t profile;      /* 1: time kernels with HIP events (ccd_kernel_stats, ccd_loop_time) */
    int32_t use_graph;    /* 1: replay each outer iteration from a captured HIP graph */
    int32_t profile_kinds; /* with profile = 1: bitmask of (1 << ccd_kernel_kind) to time, 0 = every kind.
                              Sweep kernels are timed by events bound to their own dispatch, which adds
                              nothing between kernels but about 5 us of device time per timed launch, so
                              with profile_kinds != 0 one launch in 5 of each selected sweep kind is timed
                              (the 1st, 6th, ...); the other kinds by event records around the launch,
                              which add a few microseconds between kernels each. */
    int32_t reserved[4];  /* zero */
} ccd_params;

/* Fills the reference defaults of src/pmf.h:4-26 (k=10, lambda=0.1, 5 outer,
 * 5 inner, variant 1, FP32, H first, zero_first 1, RMSE every iteration). */
void ccd_params_default(ccd_params* p);

/* Called after each outer iteration when passed to ccd_run.  Times in
 * seconds: rank_time = sweep kernels, update_time = residual-only kernels and
 * copies (both from HIP events when profile = 1, else 0), rmse_time = the
 * RMSE kernels; rmse is NaN when no test set is loaded.  Mirrors the line
 * printed at cdmf_ref.cpp:135-136 / cdmf_ocl_2.cpp:291-292. */
typedef void (*ccd_iter_cb)(void* user, int32_t oiter, double rank_time, double update_time, double rmse,
                            double rmse_time);

int32_t ccd_abi_version(void);

/* Number of visible HIP devices (0 when none); never fails. */
int32_t ccd_device_count(void);

int32_t ccd_create(int32_t device, ccd_ctx** out);
int32_t ccd_destroy(ccd_ctx* ctx);
const char* ccd_last_error(const ccd_ctx* ctx);

/* R as both compressed layouts: CSR (row_ptr[m+1], col_idx[nnz], csr_val[nnz])
 * and CSC (col_ptr[n+1], row_idx[nnz], csc_val[nnz]) of the same matrix.
 * Indices are validated (pointers monotone, ptr[last] == nnz, idx < dim). */
int32_t ccd_upload(ccd_ctx* ctx, int64_t m, int64_t n, uint64_t nnz, const uint32_t* row_ptr,
                   const uint32_t* col_idx, const float* csr_val, const uint32_t* col_ptr, const uint32_t* row_idx,
                   const float* csc_val);

/* COO test set (0-based).  nnz may be 0 (no RMSE). */
int32_t ccd_upload_test(ccd_ctx* ctx, uint64_t nnz, const uint32_t* row, const uint32_t* col, const float* val);

/* W rank-major [k][m], H rank-major [k][n] (the [t][i] indexing of MatData,
 * src/pmf_util.h:7-8, flattened).  Resets the residual to R (the reference's
 * start state, valid when W H^T = 0). */
int32_t ccd_set_factors(ccd_ctx* ctx, uint32_t k, const float* W, const float* H);

/* Runs maxiter outer iterations.  The first run after ccd_set_factors starts
 * from the residual R (the reference's start: W H^T = 0, skipping the
 * add-back in outer iteration 1).  A later run continues the same solve: the
 * residual then holds R - W H^T and every rank is added back first; such a
 * run must have zero_first = 0. */
int32_t ccd_run(ccd_ctx* ctx, const ccd_params* params, ccd_iter_cb cb, void* user);

int32_t ccd_get_factors(ccd_ctx* ctx, float* W, float* H);

/* Residual values R - W H^T on the observed entries, in the caller's CSR and
 * CSC orders (either pointer may be NULL). */
int32_t ccd_get_residual(ccd_ctx* ctx, float* csr_val, float* csc_val);

/* Test RMSE of the current factors, FP64 reduction on the device. */
int32_t ccd_rmse(ccd_ctx* ctx, double* out);

/* The objective CCD++ minimises, for the current factors and residual, in
 * FP64 on the device: out[0] = loss + lambda * reg, out[1] = loss = sum of
 * the squared residual over the observed entries, out[2] = reg =
 * sum_i |Omega_i| ||w_i||^2 + sum_j |Omega_j| ||h_j||^2.  The reference
 * computes the loss inside UpdateRating and drops it (src/cdmf_ref.cpp:27,38);
 * this is that reduction, fused over the residual the solver keeps.  Valid
 * after ccd_run (which leaves the residual exact), or after ccd_set_factors,
 * where the residual is R and W H^T is taken as 0. */
int32_t ccd_objective(ccd_ctx* ctx, float lambda, double* out);

/* Per-outer-iteration test RMSEs of the last ccd_run (count = maxiter). */
int32_t ccd_iter_rmse(ccd_ctx* ctx, double* out, int32_t max_count);

/* Kernel classes timed when profile = 1. */
enum ccd_kernel_kind {
    CCD_K_SWEEP_F = 0,      /* plain sweep over the first-swept layout */
    CCD_K_SWEEP_S = 1,      /* plain sweep over the second layout */
    CCD_K_FUSED_F = 2,      /* residual -/+ fused with the first sweep, first layout */
    CCD_K_FUSED_S = 3,      /* the same on the second layout */
    CCD_K_RESIDUAL = 4,     /* residual-only update */
    CCD_K_FINALIZE = 5,     /* split-row reductions */
    CCD_K_RMSE = 6,         /* test RMSE */
    CCD_K_COPY = 7,         /* factor copies */
    CCD_K_COUNT = 8
};

/* Totals for one kernel class over the last ccd_run: summed duration (ms),
 * launch count, and algorithmic HBM bytes per launch (the byte model of the
 * class for the uploaded matrix: index, value and pointer streams plus one
 * pass over each dense vector). */
int32_t ccd_kernel_stats(ccd_ctx* ctx, int32_t kind, double* total_ms, uint64_t* launches,
                         double* bytes_per_launch);

/* Wall time of the last ccd_run's outer loop, seconds (between the first
 * and last HIP event of the loop; 0 when profile = 0). */
int32_t ccd_loop_time(ccd_ctx* ctx, double* seconds);

/* Stream-copy calibration kernel: copies `bytes` (multiple of 16) between two
 * device buffers `reps` times; returns the mean kernel time in ms.  Used to
 * scale rocprofv3 FETCH_SIZE/WRITE_SIZE counters. */
int32_t ccd_copy_probe(ccd_ctx* ctx, uint64_t bytes, int32_t reps, double* ms_per_copy);

/* Work-plan knobs of the binned and tiled sweeps; they only change speed and
 * the order of floating-point sums.  The binned knobs apply at the next
 * ccd_upload, sigma at the next run of the tiled variant. */
typedef struct ccd_plan_params {
    uint32_t elems_per_lane;  /* target elements per lane when sizing sub-wave groups (default 4) */
    uint32_t split_threshold; /* rows longer than this are split across waves (default 8192) */
    uint32_t piece;           /* elements per piece of a split row (default 4096) */
    uint32_t unit_elems;      /* target elements per wave work unit (default 4096) */
    uint32_t sigma;           /* elements per lane of a tile of the tiled sweeps: 8, 16 or 32 (default 16; 32 runs
                                 512-thread blocks, FP32 only) */
    uint32_t blocks_per_cu;   /* persistent sweep blocks per CU: 0 = as many as registers allow (1 or 2),
                                 1 or 2 = at most that many; the staged LDS prefix is sized to match */
    uint32_t split_rows;      /* tiled sweeps, layouts whose gathered vector exceeds the LDS prefix: 1 = keep
                                 each row's elements gathered from beyond the prefix in tiles of their own,
                                 finished by a per-row finish pass (default); 0 = one tile range in row order */
} ccd_plan_params;

int32_t ccd_set_plan(ccd_ctx* ctx, const ccd_plan_params* pp);

/* ---------------------------------------------------------------- batched ---
 * Many independent factorizations (hyperparameter sweeps, per-segment
 * models) solved together by ONE persistent launch.  No counterpart in the
 * reference, which solves one matrix per process (src/main.cpp:38-64).
 *
 * Each problem is an ordinary context (ccd_upload + ccd_set_factors done)
 * with its own ccd_params (k, lambda, maxiter, maxinneriter, order,
 * zero_first; FP32 only, maxinneriter >= 1).  The batch borrows the
 * contexts: keep them alive and do not use them while the batch runs.
 * After ccd_batch_run every context holds the result of its own ccd_run
 * with the tiled sweeps on the plain (unsplit) tile layout, i.e. with
 * ccd_plan_params.split_rows = 0 (bit-identical, same sigma; a ccd_run with
 * split rows rounds differently wherever it splits a row): read it with ccd_get_factors / ccd_get_residual / ccd_rmse,
 * continue it with another batch run or a ccd_run.
 */
typedef struct ccd_batch ccd_batch;

typedef struct ccd_batch_plan {
    uint32_t sigma;       /* elements per lane of a tile: 8 or 16 (default 16) */
    uint32_t blocks;      /* persistent blocks of 256 threads, a multiple of 8; 0 = all co-resident ones */
    float work_exponent;  /* blocks are dealt to problems by (their work)^work_exponent (default 1) */
    uint32_t solo;        /* the `solo` problems of largest work (k * maxiter * (2 maxinneriter + 2) * nnz)
                             are solved first, one after another, each on the whole GPU (ccd_run of the
                             tiled sweeps, plain tile layout, same sigma: the same result); the persistent
                             launch takes the rest (default 0, at most 64).  A team of the launch is
                             confined to one XCD, so a problem with more than an eighth of the work bounds
                             the launch. */
    uint32_t deal;        /* a team's sweep over a layout with at least `deal` tiles per wave of the team
                             deals them through a counter in LDS (blocks own interleaved tiles, waves take
                             the next one when they finish one); smaller layouts are split statically
                             among the waves.  0 = the default, 16; 0xffffffff = never.  The results do
                             not depend on it. */
} ccd_batch_plan;

int32_t ccd_batch_create(int32_t device, ccd_batch** out);
int32_t ccd_batch_destroy(ccd_batch* b);
const char* ccd_batch_last_error(ccd_batch* b);
/* returns the problem's index (>= 0) or an error code */
int32_t ccd_batch_add(ccd_batch* b, ccd_ctx* ctx, const ccd_params* p);
int32_t ccd_batch_set_plan(ccd_batch* b, const ccd_batch_plan* pp);
int32_t ccd_batch_get_plan(ccd_batch* b, ccd_batch_plan* out);
/* Autotuner: times short trial launches (about 1/64 of each problem's
 * sweeps, no state change) for every sigma and work exponent and keeps the
 * fastest; then, when no problem continues an earlier run, times whole runs
 * with 0 to 3 problems solved alone (solo), then with the tile deal at 8, 32
 * and never (deal; the solo runs used the default, 16), putting every
 * problem's factors and residual back to their start after each, and keeps
 * the fastest.  *best_ms = the kept plan's trial time, or its whole-run time
 * when whole runs were timed. */
int32_t ccd_batch_tune(ccd_batch* b, double* best_ms);
/* One run of every problem: set-up fills and retiling, the persistent
 * launch, write-back.  use_graph: capture that sequence in a HIP graph on
 * first use and replay it while the plan, the problems' continued /
 * zero_first state and their device buffers stay the same (a re-upload, a
 * new k or a ccd_run that rebuilds a tile layout recaptures it). */
int32_t ccd_batch_run(ccd_batch* b, int32_t use_graph);
/* time of the last run: host time of the problems solved alone plus the
 * device time of the launch sequence (events around it) */
int32_t ccd_batch_time(ccd_batch* b, double* seconds);
/* Whole-run times (ms) the last ccd_batch_tune measured: solo = 0, 1, ...,
 * min(3, problems - 1), then deal = 8, 32, 0xffffffff with the best solo
 * count (none when it timed trials only); returns how many there are. */
int32_t ccd_batch_tune_times(ccd_batch* b, double* ms, int32_t max_count);
/* blocks given to each problem by the last run (0: solved alone); returns
 * the problem count */
int32_t ccd_batch_teams(ccd_batch* b, uint32_t* blocks, int32_t max_count);

#ifdef __cplusplus
}
#endif

#endif /* CDMF_HIP_H */
