/*
 * CPU oracle of the fold-in extension -- TEST INFRASTRUCTURE ONLY.
 *
 * Folding in a new user i (ratings r_ij, j in Omega_i) with the item
 * factors H fixed is CCD++ restricted to the row: the u-side rank-one
 * update of src/cdmf_ref.cpp:7-17 (with the item side frozen) applied to
 * one row at a time.  The reference has no fold-in; these functions define
 * the arithmetic the GPU kernels follow, and tests/test_oracle_foldin.py
 * pins them: many sweeps reach the ridge solve
 *   (H_Omega^T H_Omega + lambda |Omega_i| I) w_i = H_Omega^T r_i,
 * the features match an independent numpy formula, and the MLP gradients
 * match central finite differences.
 *
 *   or_fold_rows      FP32, per row in element order: residual init
 *                     r^ = r - sum_t w_t h_t (t ascending), then per sweep
 *                     and rank t: a = r^ + w_t h, w_t = sum a h / (lambda d +
 *                     sum h^2) (0 for an empty row), r^ = a - w_t h
 *   or_fold_features  FP64: rating-weighted mean and variance of the rated
 *                     items' H columns (2k values), log(1 + d), mean rating
 *   or_mlp_forward    FP64 MLP in -> hid -> hid -> out, GELU (erf form)
 *   or_mlp_backward   FP64 gradient of mean((Y - T)^2) w.r.t. all parameters
 *
 * Build with -ffp-contract=off (oracle/Makefile).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>

#define OR_EXPORT __attribute__((visibility("default")))

OR_EXPORT int or_fold_rows(int64_t nrows, const uint32_t* ptr, const uint32_t* idx, const float* val, uint32_t k,
                           int64_t n, const float* H, float lambda, int sweeps, float* W, int threads) {
    if (threads > 0) omp_set_num_threads(threads);
    int bad = 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(| : bad)
    for (int64_t i = 0; i < nrows; ++i) {
        const uint32_t s = ptr[i], e = ptr[i + 1], d = e - s;
        float* w = W + (size_t)i * k;
        float* r = (float*)malloc(sizeof(float) * (d ? d : 1));
        if (!r) { bad = 1; continue; }
        for (uint32_t q = 0; q < d; ++q) {
            float v = val[s + q];
            for (uint32_t t = 0; t < k; ++t) v = v - w[t] * H[(size_t)t * n + idx[s + q]];
            r[q] = v;
        }
        for (int sw = 0; sw < sweeps; ++sw) {
            for (uint32_t t = 0; t < k; ++t) {
                const float* ht = H + (size_t)t * n;
                const float wo = w[t];
                float g = 0.0f, hh = 0.0f;
                for (uint32_t q = 0; q < d; ++q) {
                    const float h = ht[idx[s + q]];
                    const float a = r[q] + wo * h;
                    r[q] = a;
                    g += a * h;
                    hh += h * h;
                }
                const float wn = d ? g / (lambda * (float)d + hh) : 0.0f;
                for (uint32_t q = 0; q < d; ++q) r[q] = r[q] - wn * ht[idx[s + q]];
                w[t] = wn;
            }
        }
        free(r);
    }
    return bad ? -1 : 0;
}

OR_EXPORT void or_fold_features(int64_t nrows, const uint32_t* ptr, const uint32_t* idx, const float* val,
                                uint32_t k, int64_t n, const float* H, float* F) {
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t i = 0; i < nrows; ++i) {
        const uint32_t s = ptr[i], e = ptr[i + 1], d = e - s;
        float* f = F + (size_t)i * (2 * k + 2);
        double S = 0.0;
        for (uint32_t q = s; q < e; ++q) S += val[q];
        for (uint32_t t = 0; t < k; ++t) {
            const float* ht = H + (size_t)t * n;
            double m = 0.0, v = 0.0;
            if (d && S != 0.0) {
                for (uint32_t q = s; q < e; ++q) m += (double)val[q] * ht[idx[q]];
                m /= S;
                for (uint32_t q = s; q < e; ++q) {
                    const double z = (double)ht[idx[q]] - m;
                    v += (double)val[q] * z * z;
                }
                v /= S;
            }
            f[t] = (float)m;
            f[k + t] = (float)v;
        }
        f[2 * k] = (float)log1p((double)d);
        f[2 * k + 1] = d ? (float)(S / d) : 0.0f;
    }
}

/* ---------------------------------------------------------------- MLP --- */
/* Parameters, flat: W1 [hid][in], b1 [hid], W2 [hid][hid], b2 [hid],
 * W3 [out][hid], b3 [out]. */

static double gelu(double x) { return 0.5 * x * (1.0 + erf(x / sqrt(2.0))); }
static double gelu_d(double x) {
    return 0.5 * (1.0 + erf(x / sqrt(2.0))) + x * exp(-0.5 * x * x) / sqrt(2.0 * 3.14159265358979323846);
}

static void layer(const double* X, int64_t nrow, int in, int out, const double* W, const double* b, double* Z) {
    for (int64_t r = 0; r < nrow; ++r)
        for (int o = 0; o < out; ++o) {
            double z = b[o];
            for (int q = 0; q < in; ++q) z += X[r * in + q] * W[(size_t)o * in + q];
            Z[r * out + o] = z;
        }
}

OR_EXPORT void or_mlp_forward(int64_t nrow, int in, int hid, int out, const double* P, const double* X, double* Y) {
    const double *W1 = P, *b1 = W1 + (size_t)hid * in, *W2 = b1 + hid, *b2 = W2 + (size_t)hid * hid,
                 *W3 = b2 + hid, *b3 = W3 + (size_t)out * hid;
    double* Z1 = (double*)malloc(sizeof(double) * nrow * hid);
    double* Z2 = (double*)malloc(sizeof(double) * nrow * hid);
    if (!Z1 || !Z2) { free(Z1); free(Z2); return; }
    layer(X, nrow, in, hid, W1, b1, Z1);
    for (int64_t q = 0; q < nrow * hid; ++q) Z1[q] = gelu(Z1[q]);
    layer(Z1, nrow, hid, hid, W2, b2, Z2);
    for (int64_t q = 0; q < nrow * hid; ++q) Z2[q] = gelu(Z2[q]);
    layer(Z2, nrow, hid, out, W3, b3, Y);
    free(Z1);
    free(Z2);
}

/* loss = mean over nrow*out of (Y - T)^2; G receives d loss / d P */
OR_EXPORT double or_mlp_backward(int64_t nrow, int in, int hid, int out, const double* P, const double* X,
                                 const double* T, double* G) {
    const double *W1 = P, *b1 = W1 + (size_t)hid * in, *W2 = b1 + hid, *b2 = W2 + (size_t)hid * hid,
                 *W3 = b2 + hid, *b3 = W3 + (size_t)out * hid;
    double *gW1 = G, *gb1 = gW1 + (size_t)hid * in, *gW2 = gb1 + hid, *gb2 = gW2 + (size_t)hid * hid,
           *gW3 = gb2 + hid, *gb3 = gW3 + (size_t)out * hid;
    const size_t np = (size_t)hid * in + hid + (size_t)hid * hid + hid + (size_t)out * hid + out;
    memset(G, 0, sizeof(double) * np);
    double *Z1 = malloc(sizeof(double) * nrow * hid), *A1 = malloc(sizeof(double) * nrow * hid),
           *Z2 = malloc(sizeof(double) * nrow * hid), *A2 = malloc(sizeof(double) * nrow * hid),
           *Y = malloc(sizeof(double) * nrow * out), *dY = malloc(sizeof(double) * nrow * out),
           *dA = malloc(sizeof(double) * nrow * hid), *dZ = malloc(sizeof(double) * nrow * hid);
    double loss = NAN;
    if (!Z1 || !A1 || !Z2 || !A2 || !Y || !dY || !dA || !dZ) goto done;
    layer(X, nrow, in, hid, W1, b1, Z1);
    for (int64_t q = 0; q < nrow * hid; ++q) A1[q] = gelu(Z1[q]);
    layer(A1, nrow, hid, hid, W2, b2, Z2);
    for (int64_t q = 0; q < nrow * hid; ++q) A2[q] = gelu(Z2[q]);
    layer(A2, nrow, hid, out, W3, b3, Y);
    loss = 0.0;
    const double scale = 1.0 / ((double)nrow * out);
    for (int64_t q = 0; q < nrow * out; ++q) {
        const double dlt = Y[q] - T[q];
        loss += dlt * dlt;
        dY[q] = 2.0 * dlt * scale;
    }
    loss *= scale;
    /* layer 3 */
    for (int64_t r = 0; r < nrow; ++r)
        for (int o = 0; o < out; ++o) {
            gb3[o] += dY[r * out + o];
            for (int q = 0; q < hid; ++q) gW3[(size_t)o * hid + q] += dY[r * out + o] * A2[r * hid + q];
        }
    for (int64_t r = 0; r < nrow; ++r)
        for (int q = 0; q < hid; ++q) {
            double a = 0.0;
            for (int o = 0; o < out; ++o) a += dY[r * out + o] * W3[(size_t)o * hid + q];
            dZ[r * hid + q] = a * gelu_d(Z2[r * hid + q]);
        }
    /* layer 2 */
    for (int64_t r = 0; r < nrow; ++r)
        for (int o = 0; o < hid; ++o) {
            gb2[o] += dZ[r * hid + o];
            for (int q = 0; q < hid; ++q) gW2[(size_t)o * hid + q] += dZ[r * hid + o] * A1[r * hid + q];
        }
    for (int64_t r = 0; r < nrow; ++r)
        for (int q = 0; q < hid; ++q) {
            double a = 0.0;
            for (int o = 0; o < hid; ++o) a += dZ[r * hid + o] * W2[(size_t)o * hid + q];
            dA[r * hid + q] = a * gelu_d(Z1[r * hid + q]);
        }
    /* layer 1 */
    for (int64_t r = 0; r < nrow; ++r)
        for (int o = 0; o < hid; ++o) {
            gb1[o] += dA[r * hid + o];
            for (int q = 0; q < in; ++q) gW1[(size_t)o * in + q] += dA[r * hid + o] * X[r * in + q];
        }
done:
    free(Z1); free(A1); free(Z2); free(A2); free(Y); free(dY); free(dA); free(dZ);
    return loss;
}
