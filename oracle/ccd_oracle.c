/*
 * CPU oracle for the CCD++ hot path -- TEST INFRASTRUCTURE ONLY.
 *
 * This file restates, in plain C with OpenMP, the arithmetic of the
 * reference solver so that the HIP path can be checked against it.  Only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
 * the library built from it; the product never links it.
 *
 * Functions and the reference code each one follows:
 *   or_rank_one_update   src/cdmf_ref.cpp:7-17    (RankOneUpdate, h seeded with lambda*|Omega_j|)
 *   or_update_rating     src/cdmf_ref.cpp:19-44   (UpdateRating, loss discarded at its call sites)
 *   or_ccd               src/cdmf_ref.cpp:47-140  (cdmf_ref driver: H := 0, oiter/t/inner loops,
 *                                                 add-back skipped at oiter 1, v-sweep then u-sweep)
 *   or_rmse              src/tools.cpp:543-582    (calculate_rmse_directly, float products summed in
 *                                                 double, NaN terms skipped but counted)
 *   or_initial_col       src/tools.cpp:349-357    (initial_col: srand(0), i-major draw order)
 *   or_dual_spmv         src/detail/kcode/csr5_spmv.cl:134-147 (dual outputs y_t = A x, y_b = pattern x^2)
 *   or_ccd_f64           FP64 twin of or_ccd (precision study; no reference counterpart)
 *   or_ccd_bf16          or_ccd with every stored W/H value rounded to bfloat16 (no reference counterpart)
 *   or_objective         the objective CCD++ minimises, in FP64 (cdmf_ref.cpp:27,38 computes sum r^2 and drops it)
 *   or_ccd_ranks         ranks [t0, t1) of one outer iteration of or_ccd, in place (CPU baseline)
 *
 * Train RMSE: every solver here can also report, per outer iteration, the
 * training RMSE sqrt(sum rhat^2 / nnz) of the residual R - W H^T left after
 * that iteration, summed in double in CSR order.  The reference computes
 * sum rhat^2 inside UpdateRating (cdmf_ref.cpp:27,38) and discards it; this
 * is that quantity for the whole matrix, taken where each outer iteration
 * ends.
 *
 * Pinning: or_ccd / or_rmse / or_initial_col are checked bit-for-bit against
 * fixtures written by the compiled reference (oracle/Makefile, tests/golden/).
 * or_ccd_ranks equals or_ccd bit-for-bit when its rank slices are chained.
 * or_ccd_bf16 equals an independent numpy restatement with the same rounding
 * points bit-for-bit, and, built with OR_BF16_IDENTITY (rounding replaced
 * by the identity), equals or_ccd bit-for-bit.  or_ccd_f64, or_dual_spmv,
 * or_objective and the train RMSE are pinned by mathematical checks in
 * tests/test_oracle.py (closed-form ridge solves, a monotone objective, the
 * transpose symmetry, scipy SpMV, a numpy recomputation).
 *
 * Build with -ffp-contract=off so that products and sums round exactly as in
 * the reference build (plain x86-64 SSE, no FMA).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>

#define OR_EXPORT __attribute__((visibility("default")))

/* ---------------------------------------------------------------------- */
/* FP32 restatement                                                         */
/* ---------------------------------------------------------------------- */

/* cdmf_ref.cpp:7-17 -- closed-form rank-one coordinate update for one column
 * j of a compressed layout: g = sum u_i r_ij, h = lambda_n + sum u_i^2, in
 * index order; an empty column yields 0. */
static inline float or_rank_one_update(const uint32_t* ptr, const uint32_t* idx, const float* val,
                                       int64_t j, const float* u, float lambda_n) {
    float g = 0.0f, h = lambda_n;
    if (ptr[j + 1] == ptr[j]) return 0.0f;
    for (uint32_t e = ptr[j]; e < ptr[j + 1]; ++e) {
        uint32_t i = idx[e];
        g += u[i] * val[e];
        h += u[i] * u[i];
    }
    return g / h;
}

/* cdmf_ref.cpp:19-44 -- val[e] +/-= Wt[idx[e]] * Ht[c] over every column c of
 * one compressed layout.  The loss it also accumulates is unused by callers. */
static void or_update_rating(int64_t cols, const uint32_t* ptr, const uint32_t* idx, float* val,
                             const float* Wt, const float* Ht, int add) {
    if (add) {
#pragma omp parallel for schedule(dynamic, 500)
        for (int64_t c = 0; c < cols; ++c) {
            float Htc = Ht[c];
            for (uint32_t e = ptr[c]; e < ptr[c + 1]; ++e) val[e] += Wt[idx[e]] * Htc;
        }
    } else {
#pragma omp parallel for schedule(dynamic, 500)
        for (int64_t c = 0; c < cols; ++c) {
            float Htc = Ht[c];
            for (uint32_t e = ptr[c]; e < ptr[c + 1]; ++e) val[e] -= Wt[idx[e]] * Htc;
        }
    }
}

/* tools.cpp:543-582 -- test RMSE: pred = sum_t W[t][i]*H[t][j] (float products
 * accumulated in double), squared errors summed in double in entry order,
 * NaN terms skipped but counted in the denominator.  W is [k][m], H is [k][n]. */
OR_EXPORT double or_rmse(const float* W, const float* H, int64_t m, int64_t n, uint32_t k,
                         uint64_t nnz, const uint32_t* trow, const uint32_t* tcol, const float* tval) {
    double rmse = 0.0;
    uint64_t num = 0;
    for (uint64_t e = 0; e < nnz; ++e) {
        uint32_t i = trow[e], j = tcol[e];
        double v = tval[e];
        double pred = 0.0;
        for (uint32_t t = 0; t < k; ++t) {
            float p = W[(size_t)t * m + i] * H[(size_t)t * n + j];
            pred += p;
        }
        double tmp = (pred - v) * (pred - v);
        if (!isnan(tmp)) rmse += tmp;
        ++num;
    }
    if (num == 0) return NAN;
    return sqrt(rmse / (double)num);
}

/* Train RMSE of a residual array: sqrt(sum r^2 / nnz), summed in double in
 * array order; NaN for an empty matrix. */
static double or_train_rmse(const float* res, uint64_t nnz) {
    double s = 0.0;
    for (uint64_t e = 0; e < nnz; ++e) s += (double)res[e] * (double)res[e];
    return nnz ? sqrt(s / (double)nnz) : NAN;
}

static double or_train_rmse_f64(const double* res, uint64_t nnz) {
    double s = 0.0;
    for (uint64_t e = 0; e < nnz; ++e) s += res[e] * res[e];
    return nnz ? sqrt(s / (double)nnz) : NAN;
}

/* tools.cpp:349-357 -- X[j][i] = 0.1*(rand()/RAND_MAX) + 0.001 with srand(0)
 * before the draws, i-major order.  X is [k][n] rank-major. */
OR_EXPORT void or_initial_col(float* X, uint32_t k, uint32_t n) {
    srand(0);
    for (uint32_t i = 0; i < n; ++i)
        for (uint32_t j = 0; j < k; ++j)
            X[(size_t)j * n + i] = 0.1f * ((float)rand() / (float)RAND_MAX) + 0.001f;
}

/*
 * cdmf_ref.cpp:47-140 -- the whole CCD++ loop, generalised over which layout
 * is swept first.
 *
 *   order 0 (the reference): v-sweep over the CSC copy first, then the u-sweep
 *            over the CSR copy; the solver zeroes H at entry.
 *   order 1: u-sweep first.  This is exactly order 0 applied to R^T with the
 *            roles of W and H exchanged (the shallow transpose of
 *            pmf_util.h:38-53); it zeroes W at entry.
 *
 * zero_first:  1 reproduces the reference's zeroing of the first-swept factor
 *              (cdmf_ref.cpp:57-61); 0 keeps the caller's values (the caller
 *              then guarantees that their product is zero, e.g. W = 0).
 * R's value arrays are updated in place and hold the residual on return,
 * as in the reference.  rmse_out and train_out (either may be NULL) receive
 * maxiter test and train RMSEs.
 */
OR_EXPORT int or_ccd(int64_t m, int64_t n,
                     const uint32_t* csr_ptr, const uint32_t* csr_idx, float* csr_val,
                     const uint32_t* csc_ptr, const uint32_t* csc_idx, float* csc_val,
                     float* W, float* H, uint32_t k, float lambda, int maxiter, int maxinner,
                     int order, int zero_first, int threads,
                     uint64_t nnz_test, const uint32_t* trow, const uint32_t* tcol, const float* tval,
                     double* rmse_out, double* train_out) {
    /* "first" layout F is swept first and produces factor XF; "second" layout S
     * produces XS.  order 0: F = CSC (dim n, gathers W), S = CSR (dim m). */
    const int64_t dF = order ? m : n, dS = order ? n : m;
    const uint32_t* pF = order ? csr_ptr : csc_ptr;
    const uint32_t* iF = order ? csr_idx : csc_idx;
    float* vF = order ? csr_val : csc_val;
    const uint32_t* pS = order ? csc_ptr : csr_ptr;
    const uint32_t* iS = order ? csc_idx : csr_idx;
    float* vS = order ? csc_val : csr_val;
    float* XF = order ? W : H; /* output of the first sweep, [k][dF] */
    float* XS = order ? H : W; /* output of the second sweep, [k][dS] */

    int old_threads = omp_get_max_threads();
    if (threads > 0) omp_set_num_threads(threads);

    if (zero_first) memset(XF, 0, sizeof(float) * (size_t)k * (size_t)dF);

    float* u = (float*)malloc(sizeof(float) * (size_t)(dS > 0 ? dS : 1)); /* gathered by the first sweep */
    float* v = (float*)malloc(sizeof(float) * (size_t)(dF > 0 ? dF : 1));
    if (!u || !v) { free(u); free(v); return -1; }

    for (int oiter = 1; oiter <= maxiter; ++oiter) {
        for (uint32_t t = 0; t < k; ++t) {
            float* St = XS + (size_t)t * dS;
            float* Ft = XF + (size_t)t * dF;
#pragma omp parallel for
            for (int64_t i = 0; i < dS; ++i) u[i] = St[i];
#pragma omp parallel for
            for (int64_t i = 0; i < dF; ++i) v[i] = Ft[i];

            if (oiter > 1) {
                or_update_rating(dF, pF, iF, vF, St, Ft, 1);
                or_update_rating(dS, pS, iS, vS, Ft, St, 1);
            }
            for (int it = 1; it <= maxinner; ++it) {
#pragma omp parallel for schedule(dynamic, 500)
                for (int64_t c = 0; c < dF; ++c)
                    v[c] = or_rank_one_update(pF, iF, vF, c, u, lambda * (float)(pF[c + 1] - pF[c]));
#pragma omp parallel for schedule(dynamic, 500)
                for (int64_t c = 0; c < dS; ++c)
                    u[c] = or_rank_one_update(pS, iS, vS, c, v, lambda * (float)(pS[c + 1] - pS[c]));
            }
#pragma omp parallel for
            for (int64_t i = 0; i < dS; ++i) St[i] = u[i];
#pragma omp parallel for
            for (int64_t i = 0; i < dF; ++i) Ft[i] = v[i];
            or_update_rating(dF, pF, iF, vF, u, v, 0);
            or_update_rating(dS, pS, iS, vS, v, u, 0);
        }
        if (rmse_out) rmse_out[oiter - 1] = or_rmse(W, H, m, n, k, nnz_test, trow, tcol, tval);
        if (train_out) train_out[oiter - 1] = or_train_rmse(csr_val, csr_ptr[m]);
    }
    free(u);
    free(v);
    omp_set_num_threads(old_threads);
    return 0;
}

/*
 * A bounded slice of or_ccd for timing: ranks [t0, t1) of one outer iteration
 * `oiter`, with the same loops and schedule.  Used only as the CPU baseline.
 */
OR_EXPORT int or_ccd_ranks(int64_t m, int64_t n,
                           const uint32_t* csr_ptr, const uint32_t* csr_idx, float* csr_val,
                           const uint32_t* csc_ptr, const uint32_t* csc_idx, float* csc_val,
                           float* W, float* H, uint32_t t0, uint32_t t1, float lambda, int oiter,
                           int maxinner, int threads) {
    float* u = (float*)malloc(sizeof(float) * (size_t)m);
    float* v = (float*)malloc(sizeof(float) * (size_t)n);
    if (!u || !v) { free(u); free(v); return -1; }
    int old_threads = omp_get_max_threads();
    if (threads > 0) omp_set_num_threads(threads);
    for (uint32_t t = t0; t < t1; ++t) {
        float* Wt = W + (size_t)t * m;
        float* Ht = H + (size_t)t * n;
#pragma omp parallel for
        for (int64_t i = 0; i < m; ++i) u[i] = Wt[i];
#pragma omp parallel for
        for (int64_t i = 0; i < n; ++i) v[i] = Ht[i];
        if (oiter > 1) {
            or_update_rating(n, csc_ptr, csc_idx, csc_val, Wt, Ht, 1);
            or_update_rating(m, csr_ptr, csr_idx, csr_val, Ht, Wt, 1);
        }
        for (int it = 1; it <= maxinner; ++it) {
#pragma omp parallel for schedule(dynamic, 500)
            for (int64_t c = 0; c < n; ++c)
                v[c] = or_rank_one_update(csc_ptr, csc_idx, csc_val, c, u, lambda * (float)(csc_ptr[c + 1] - csc_ptr[c]));
#pragma omp parallel for schedule(dynamic, 500)
            for (int64_t c = 0; c < m; ++c)
                u[c] = or_rank_one_update(csr_ptr, csr_idx, csr_val, c, v, lambda * (float)(csr_ptr[c + 1] - csr_ptr[c]));
        }
#pragma omp parallel for
        for (int64_t i = 0; i < m; ++i) Wt[i] = u[i];
#pragma omp parallel for
        for (int64_t i = 0; i < n; ++i) Ht[i] = v[i];
        or_update_rating(n, csc_ptr, csc_idx, csc_val, u, v, 0);
        or_update_rating(m, csr_ptr, csr_idx, csr_val, v, u, 0);
    }
    omp_set_num_threads(old_threads);
    free(u);
    free(v);
    return 0;
}

/* ---------------------------------------------------------------------- */
/* BF16-storage twin: every value written to W, H, u or v is rounded to the   */
/* nearest bfloat16 (ties to even); arithmetic stays FP32.                    */
/* ---------------------------------------------------------------------- */

#ifdef OR_BF16_IDENTITY
/* Check build (oracle/Makefile, libccd_oracle_bf16id.so): no rounding, so
 * or_ccd_bf16 must reproduce or_ccd bit for bit. */
static inline float or_bf16_round(float x) { return x; }
#else
static inline float or_bf16_round(float x) {
    uint32_t b;
    memcpy(&b, &x, 4);
    if ((b & 0x7f800000u) == 0x7f800000u) { /* inf / nan: keep, quiet nan */
        if (b & 0x007fffffu) b |= 0x00400000u;
        b &= 0xffff0000u;
    } else {
        uint32_t lsb = (b >> 16) & 1u;
        b += 0x7fffu + lsb;
        b &= 0xffff0000u;
    }
    float y;
    memcpy(&y, &b, 4);
    return y;
}
#endif

OR_EXPORT float or_bf16(float x) { return or_bf16_round(x); }

OR_EXPORT int or_ccd_bf16(int64_t m, int64_t n,
                          const uint32_t* csr_ptr, const uint32_t* csr_idx, float* csr_val,
                          const uint32_t* csc_ptr, const uint32_t* csc_idx, float* csc_val,
                          float* W, float* H, uint32_t k, float lambda, int maxiter, int maxinner,
                          int order, int zero_first, int threads,
                          uint64_t nnz_test, const uint32_t* trow, const uint32_t* tcol, const float* tval,
                          double* rmse_out, double* train_out) {
    const int64_t dF = order ? m : n, dS = order ? n : m;
    const uint32_t* pF = order ? csr_ptr : csc_ptr;
    const uint32_t* iF = order ? csr_idx : csc_idx;
    float* vF = order ? csr_val : csc_val;
    const uint32_t* pS = order ? csc_ptr : csr_ptr;
    const uint32_t* iS = order ? csc_idx : csr_idx;
    float* vS = order ? csc_val : csr_val;
    float* XF = order ? W : H;
    float* XS = order ? H : W;
    int old_threads = omp_get_max_threads();
    if (threads > 0) omp_set_num_threads(threads);
    if (zero_first) memset(XF, 0, sizeof(float) * (size_t)k * (size_t)dF);
    /* the stored factors are bf16 from the start */
    for (size_t e = 0; e < (size_t)k * (size_t)dF; ++e) XF[e] = or_bf16_round(XF[e]);
    for (size_t e = 0; e < (size_t)k * (size_t)dS; ++e) XS[e] = or_bf16_round(XS[e]);
    float* u = (float*)malloc(sizeof(float) * (size_t)(dS > 0 ? dS : 1));
    float* v = (float*)malloc(sizeof(float) * (size_t)(dF > 0 ? dF : 1));
    if (!u || !v) { free(u); free(v); return -1; }
    for (int oiter = 1; oiter <= maxiter; ++oiter) {
        for (uint32_t t = 0; t < k; ++t) {
            float* St = XS + (size_t)t * dS;
            float* Ft = XF + (size_t)t * dF;
            for (int64_t i = 0; i < dS; ++i) u[i] = St[i];
            for (int64_t i = 0; i < dF; ++i) v[i] = Ft[i];
            if (oiter > 1) {
                or_update_rating(dF, pF, iF, vF, St, Ft, 1);
                or_update_rating(dS, pS, iS, vS, Ft, St, 1);
            }
            for (int it = 1; it <= maxinner; ++it) {
#pragma omp parallel for schedule(dynamic, 500)
                for (int64_t c = 0; c < dF; ++c)
                    v[c] = or_bf16_round(or_rank_one_update(pF, iF, vF, c, u, lambda * (float)(pF[c + 1] - pF[c])));
#pragma omp parallel for schedule(dynamic, 500)
                for (int64_t c = 0; c < dS; ++c)
                    u[c] = or_bf16_round(or_rank_one_update(pS, iS, vS, c, v, lambda * (float)(pS[c + 1] - pS[c])));
            }
            for (int64_t i = 0; i < dS; ++i) St[i] = u[i];
            for (int64_t i = 0; i < dF; ++i) Ft[i] = v[i];
            or_update_rating(dF, pF, iF, vF, u, v, 0);
            or_update_rating(dS, pS, iS, vS, v, u, 0);
        }
        if (rmse_out) rmse_out[oiter - 1] = or_rmse(W, H, m, n, k, nnz_test, trow, tcol, tval);
        if (train_out) train_out[oiter - 1] = or_train_rmse(csr_val, csr_ptr[m]);
    }
    free(u);
    free(v);
    omp_set_num_threads(old_threads);
    return 0;
}

/* ---------------------------------------------------------------------- */
/* FP64 twin (precision study).  Values, residual and factors in double.     */
/* ---------------------------------------------------------------------- */

static inline double or_rank_one_update_f64(const uint32_t* ptr, const uint32_t* idx, const double* val,
                                            int64_t j, const double* u, double lambda_n) {
    double g = 0.0, h = lambda_n;
    if (ptr[j + 1] == ptr[j]) return 0.0;
    for (uint32_t e = ptr[j]; e < ptr[j + 1]; ++e) {
        uint32_t i = idx[e];
        g += u[i] * val[e];
        h += u[i] * u[i];
    }
    return g / h;
}

static void or_update_rating_f64(int64_t cols, const uint32_t* ptr, const uint32_t* idx, double* val,
                                 const double* Wt, const double* Ht, int add) {
    const double s = add ? 1.0 : -1.0;
#pragma omp parallel for schedule(dynamic, 500)
    for (int64_t c = 0; c < cols; ++c) {
        double Htc = Ht[c];
        for (uint32_t e = ptr[c]; e < ptr[c + 1]; ++e) val[e] += s * (Wt[idx[e]] * Htc);
    }
}

OR_EXPORT double or_rmse_f64(const double* W, const double* H, int64_t m, int64_t n, uint32_t k,
                             uint64_t nnz, const uint32_t* trow, const uint32_t* tcol, const float* tval) {
    double rmse = 0.0;
    uint64_t num = 0;
    for (uint64_t e = 0; e < nnz; ++e) {
        uint32_t i = trow[e], j = tcol[e];
        double pred = 0.0;
        for (uint32_t t = 0; t < k; ++t) pred += W[(size_t)t * m + i] * H[(size_t)t * n + j];
        double tmp = (pred - (double)tval[e]) * (pred - (double)tval[e]);
        if (!isnan(tmp)) rmse += tmp;
        ++num;
    }
    if (num == 0) return NAN;
    return sqrt(rmse / (double)num);
}

OR_EXPORT int or_ccd_f64(int64_t m, int64_t n,
                         const uint32_t* csr_ptr, const uint32_t* csr_idx, double* csr_val,
                         const uint32_t* csc_ptr, const uint32_t* csc_idx, double* csc_val,
                         double* W, double* H, uint32_t k, double lambda, int maxiter, int maxinner,
                         int order, int zero_first, int threads,
                         uint64_t nnz_test, const uint32_t* trow, const uint32_t* tcol, const float* tval,
                         double* rmse_out, double* train_out) {
    const int64_t dF = order ? m : n, dS = order ? n : m;
    const uint32_t* pF = order ? csr_ptr : csc_ptr;
    const uint32_t* iF = order ? csr_idx : csc_idx;
    double* vF = order ? csr_val : csc_val;
    const uint32_t* pS = order ? csc_ptr : csr_ptr;
    const uint32_t* iS = order ? csc_idx : csr_idx;
    double* vS = order ? csc_val : csr_val;
    double* XF = order ? W : H;
    double* XS = order ? H : W;
    int old_threads = omp_get_max_threads();
    if (threads > 0) omp_set_num_threads(threads);
    if (zero_first) memset(XF, 0, sizeof(double) * (size_t)k * (size_t)dF);
    double* u = (double*)malloc(sizeof(double) * (size_t)(dS > 0 ? dS : 1));
    double* v = (double*)malloc(sizeof(double) * (size_t)(dF > 0 ? dF : 1));
    if (!u || !v) { free(u); free(v); return -1; }
    for (int oiter = 1; oiter <= maxiter; ++oiter) {
        for (uint32_t t = 0; t < k; ++t) {
            double* St = XS + (size_t)t * dS;
            double* Ft = XF + (size_t)t * dF;
            memcpy(u, St, sizeof(double) * (size_t)dS);
            memcpy(v, Ft, sizeof(double) * (size_t)dF);
            if (oiter > 1) {
                or_update_rating_f64(dF, pF, iF, vF, St, Ft, 1);
                or_update_rating_f64(dS, pS, iS, vS, Ft, St, 1);
            }
            for (int it = 1; it <= maxinner; ++it) {
#pragma omp parallel for schedule(dynamic, 500)
                for (int64_t c = 0; c < dF; ++c)
                    v[c] = or_rank_one_update_f64(pF, iF, vF, c, u, lambda * (double)(pF[c + 1] - pF[c]));
#pragma omp parallel for schedule(dynamic, 500)
                for (int64_t c = 0; c < dS; ++c)
                    u[c] = or_rank_one_update_f64(pS, iS, vS, c, v, lambda * (double)(pS[c + 1] - pS[c]));
            }
            memcpy(St, u, sizeof(double) * (size_t)dS);
            memcpy(Ft, v, sizeof(double) * (size_t)dF);
            or_update_rating_f64(dF, pF, iF, vF, u, v, 0);
            or_update_rating_f64(dS, pS, iS, vS, v, u, 0);
        }
        if (rmse_out) rmse_out[oiter - 1] = or_rmse_f64(W, H, m, n, k, nnz_test, trow, tcol, tval);
        if (train_out) train_out[oiter - 1] = or_train_rmse_f64(csr_val, csr_ptr[m]);
    }
    free(u);
    free(v);
    omp_set_num_threads(old_threads);
    return 0;
}

/* ---------------------------------------------------------------------- */
/* Component restatements                                                   */
/* ---------------------------------------------------------------------- */

/* csr5_spmv.cl:134-147 -- the modified CSR5 SpMV returns two vectors per row:
 * y_t = sum val * x[col] and y_b = sum x[col]^2 (alpha is ignored).  Serial
 * sums per row in index order, accumulated in double to serve as a
 * tolerance reference for the tiled kernel. */
OR_EXPORT void or_dual_spmv(int64_t rows, const uint32_t* ptr, const uint32_t* idx, const float* val,
                            const float* x, double* yt, double* yb) {
#pragma omp parallel for schedule(dynamic, 500)
    for (int64_t r = 0; r < rows; ++r) {
        double st = 0.0, sb = 0.0;
        for (uint32_t e = ptr[r]; e < ptr[r + 1]; ++e) {
            double xv = x[idx[e]];
            st += (double)val[e] * xv;
            sb += xv * xv;
        }
        yt[r] = st;
        yb[r] = sb;
    }
}

/* One rank-one sweep over a compressed layout with the reference's formula
 * (cdmf_ref.cpp:7-17), FP64 sums: y_r = g_r / (lambda*|Omega_r| + h_r), 0 if empty. */
OR_EXPORT void or_sweep_f64(int64_t rows, const uint32_t* ptr, const uint32_t* idx, const float* val,
                            const float* x, float lambda, double* y) {
#pragma omp parallel for schedule(dynamic, 500)
    for (int64_t r = 0; r < rows; ++r) {
        if (ptr[r + 1] == ptr[r]) { y[r] = 0.0; continue; }
        double g = 0.0, h = 0.0;
        for (uint32_t e = ptr[r]; e < ptr[r + 1]; ++e) {
            double xv = x[idx[e]];
            g += xv * (double)val[e];
            h += xv * xv;
        }
        y[r] = g / ((double)lambda * (double)(ptr[r + 1] - ptr[r]) + h);
    }
}

/* Objective of CCD++ on the observed entries, in FP64:
 *   f = sum_{(i,j)} rhat_ij^2 + lambda*(sum_i |Omega_i| ||w_i||^2 + sum_j |Omega_j| ||h_j||^2)
 * rhat is taken from a CSR residual array. */
OR_EXPORT double or_objective(int64_t m, int64_t n, const uint32_t* csr_ptr, const float* csr_res,
                              const uint32_t* csc_ptr, const float* W, const float* H, uint32_t k,
                              double lambda) {
    double loss = 0.0, reg = 0.0;
    for (int64_t i = 0; i < m; ++i)
        for (uint32_t e = csr_ptr[i]; e < csr_ptr[i + 1]; ++e) loss += (double)csr_res[e] * (double)csr_res[e];
    for (int64_t i = 0; i < m; ++i) {
        double d = (double)(csr_ptr[i + 1] - csr_ptr[i]), s = 0.0;
        for (uint32_t t = 0; t < k; ++t) s += (double)W[(size_t)t * m + i] * (double)W[(size_t)t * m + i];
        reg += d * s;
    }
    for (int64_t j = 0; j < n; ++j) {
        double d = (double)(csc_ptr[j + 1] - csc_ptr[j]), s = 0.0;
        for (uint32_t t = 0; t < k; ++t) s += (double)H[(size_t)t * n + j] * (double)H[(size_t)t * n + j];
        reg += d * s;
    }
    return loss + lambda * reg;
}

/* Residual recomputed from scratch in FP64 on a CSR layout:
 * out[e] = r_ij - sum_t W[t][i] H[t][j]. */
OR_EXPORT void or_residual_f64(int64_t m, int64_t n, const uint32_t* csr_ptr, const uint32_t* csr_idx,
                               const float* csr_orig, const float* W, const float* H, uint32_t k,
                               double* out) {
#pragma omp parallel for schedule(dynamic, 500)
    for (int64_t i = 0; i < m; ++i)
        for (uint32_t e = csr_ptr[i]; e < csr_ptr[i + 1]; ++e) {
            uint32_t j = csr_idx[e];
            double p = 0.0;
            for (uint32_t t = 0; t < k; ++t) p += (double)W[(size_t)t * m + i] * (double)H[(size_t)t * n + j];
            out[e] = (double)csr_orig[e] - p;
        }
}
