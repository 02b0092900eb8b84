// C ABI of the fold-in extension (include/cdmf_foldin.h): the warm-start MLP
// (training and inference on the device with kernels/mlp_kernels.hip) and
// the fold-in of new rows (kernels/foldin_kernels.hip).
#include "cdmf_foldin.h"
#include "ccd_kernels.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <numeric>
#include <string>
#include <vector>

namespace {

template <class T>
hipError_t dalloc(T** p, size_t count) {
    *p = nullptr;
    return hipMalloc(reinterpret_cast<void**>(p), sizeof(T) * (count ? count : 1));
}

void dfree(void* p) {
    if (p) (void)hipFree(p);
}

// Grow-only device buffer.
template <class T>
hipError_t ensure(T** p, size_t* cap, size_t count) {
    if (*p && *cap >= count) return hipSuccess;
    dfree(*p);
    *p = nullptr;
    *cap = 0;
    const hipError_t e = dalloc(p, count);
    if (e == hipSuccess) *cap = count;
    return e;
}

uint64_t splitmix(uint64_t& s) {
    uint64_t z = (s += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

double uniform01(uint64_t& s) { return (double)(splitmix(s) >> 11) * (1.0 / 9007199254740992.0); }

}  // namespace

#define OBJ_TRY(o, expr)                                                                                  \
    do {                                                                                                  \
        hipError_t e_ = (expr);                                                                           \
        if (e_ != hipSuccess) {                                                                           \
            (o)->err = std::string(#expr) + ": " + hipGetErrorString(e_);                                 \
            return CCD_ERR_HIP;                                                                           \
        }                                                                                                 \
    } while (0)

static int obj_err(std::string& err, int code, const std::string& msg) {
    err = msg;
    return code;
}

// ------------------------------------------------------------------ MLP ---
struct ccd_mlp {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    uint32_t in = 0, hid = 0, out = 0;
    uint64_t np = 0;
    size_t oW1 = 0, ob1 = 0, oW2 = 0, ob2 = 0, oW3 = 0, ob3 = 0;
    float* P = nullptr;  // FP32 master parameters
    uint16_t* P16 = nullptr;
    bool p16_valid = false;
    // packed bfloat16 weights of the matrix-core inference path (mlp_mfma.hip)
    uint16_t* pk = nullptr;
    float* pkb = nullptr;
    uint32_t kin = 0, nout = 0;
    bool pk_valid = false;
    float *G = nullptr, *mom = nullptr, *vel = nullptr;
    uint64_t step = 0;
    // affine maps: inputs x_n = (x - in_sub) in_mul; outputs y = (y_n - out_sub) out_mul;
    // training targets t_n = (t - tgt_sub) tgt_mul
    float *in_sub = nullptr, *in_mul = nullptr, *out_sub = nullptr, *out_mul = nullptr, *tgt_sub = nullptr,
          *tgt_mul = nullptr;
    bool has_norm = false;
    // workspace for up to cap rows
    size_t cap = 0;
    float *Xn = nullptr, *Z1 = nullptr, *A1 = nullptr, *Z2 = nullptr, *A2 = nullptr, *Yn = nullptr, *Tn = nullptr,
          *dY = nullptr, *dZ2 = nullptr, *dZ1 = nullptr;
    double *part = nullptr, *loss = nullptr;
    size_t cap_loss = 0;
    float* gpart = nullptr;  // split-K partials of the weight-gradient products
    uint64_t gpart_cap = 0;
    // training set
    float *X = nullptr, *T = nullptr;
    uint32_t* perm = nullptr;
    size_t cap_x = 0, cap_t = 0, cap_perm = 0;
};

namespace {

int mlp_workspace(ccd_mlp* m, size_t rows) {
    if (rows <= m->cap) return CCD_OK;
    float** bufs_h[] = {&m->Z1, &m->A1, &m->Z2, &m->A2, &m->dZ2, &m->dZ1};
    for (float** b : bufs_h) {
        dfree(*b);
        OBJ_TRY(m, dalloc(b, rows * m->hid));
    }
    float** bufs_o[] = {&m->Yn, &m->Tn, &m->dY};
    for (float** b : bufs_o) {
        dfree(*b);
        OBJ_TRY(m, dalloc(b, rows * m->out));
    }
    dfree(m->Xn);
    OBJ_TRY(m, dalloc(&m->Xn, rows * m->in));
    m->cap = rows;
    return CCD_OK;
}

const void* wptr(const ccd_mlp* m, bool b16, size_t off) {
    return b16 ? static_cast<const void*>(m->P16 + off) : static_cast<const void*>(m->P + off);
}

// Yn = MLP(Xn) for n <= cap rows on stream s (workspace Z1..A2 filled).
int mlp_forward_ws(ccd_mlp* m, uint32_t n, bool b16, hipStream_t s) {
    ccd::GemmArgs g{};
    g.A = m->Xn;
    g.sai = m->in, g.sap = 1, g.sbp = 1, g.sbj = m->in;
    g.M = n, g.N = m->hid, g.K = m->in, g.ldc = m->hid;
    g.bias = wptr(m, b16, m->ob1), g.C = m->Z1, g.C2 = m->A1;
    OBJ_TRY(m, ccd::launch_gemm(g, wptr(m, b16, m->oW1), b16, ccd::GEMM_BIAS_GELU, s));
    g.A = m->A1;
    g.sai = m->hid, g.sbj = m->hid, g.K = m->hid;
    g.bias = wptr(m, b16, m->ob2), g.C = m->Z2, g.C2 = m->A2;
    OBJ_TRY(m, ccd::launch_gemm(g, wptr(m, b16, m->oW2), b16, ccd::GEMM_BIAS_GELU, s));
    g.A = m->A2;
    g.N = m->out, g.ldc = m->out;
    g.bias = wptr(m, b16, m->ob3), g.C = m->Yn, g.C2 = nullptr;
    OBJ_TRY(m, ccd::launch_gemm(g, wptr(m, b16, m->oW3), b16, ccd::GEMM_STORE, s));
    return CCD_OK;
}

// Gradient of mean((Yn - Tn)^2) into G, loss into *loss_dev (after mlp_forward_ws, FP32).
int mlp_backward_ws(ccd_mlp* m, uint32_t n, double* loss_dev, hipStream_t s) {
    OBJ_TRY(m, ccd::launch_mse(m->Yn, m->Tn, (uint64_t)n * m->out, m->dY, m->part, loss_dev, s));
    ccd::GemmArgs g{};
    // gW3 [out][hid] = dY^T A2
    g.A = m->dY, g.sai = 1, g.sap = m->out;
    g.sbp = m->hid, g.sbj = 1;
    g.M = m->out, g.N = m->hid, g.K = n, g.ldc = m->hid, g.C = m->G + m->oW3;
    g.part = m->gpart, g.part_cap = m->gpart_cap;
    OBJ_TRY(m, ccd::launch_gemm(g, m->A2, false, ccd::GEMM_STORE, s));
    OBJ_TRY(m, ccd::launch_colsum(m->dY, n, m->out, m->G + m->ob3, s));
    // dZ2 = (dY W3) * gelu'(Z2)
    g = ccd::GemmArgs{};
    g.A = m->dY, g.sai = m->out, g.sap = 1, g.sbp = m->hid, g.sbj = 1;
    g.M = n, g.N = m->hid, g.K = m->out, g.ldc = m->hid, g.C = m->dZ2, g.aux = m->Z2;
    OBJ_TRY(m, ccd::launch_gemm(g, m->P + m->oW3, false, ccd::GEMM_GELU_GRAD, s));
    // gW2 [hid][hid] = dZ2^T A1
    g = ccd::GemmArgs{};
    g.A = m->dZ2, g.sai = 1, g.sap = m->hid, g.sbp = m->hid, g.sbj = 1;
    g.M = m->hid, g.N = m->hid, g.K = n, g.ldc = m->hid, g.C = m->G + m->oW2;
    g.part = m->gpart, g.part_cap = m->gpart_cap;
    OBJ_TRY(m, ccd::launch_gemm(g, m->A1, false, ccd::GEMM_STORE, s));
    OBJ_TRY(m, ccd::launch_colsum(m->dZ2, n, m->hid, m->G + m->ob2, s));
    // dZ1 = (dZ2 W2) * gelu'(Z1)
    g = ccd::GemmArgs{};
    g.A = m->dZ2, g.sai = m->hid, g.sap = 1, g.sbp = m->hid, g.sbj = 1;
    g.M = n, g.N = m->hid, g.K = m->hid, g.ldc = m->hid, g.C = m->dZ1, g.aux = m->Z1;
    OBJ_TRY(m, ccd::launch_gemm(g, m->P + m->oW2, false, ccd::GEMM_GELU_GRAD, s));
    // gW1 [hid][in] = dZ1^T Xn
    g = ccd::GemmArgs{};
    g.A = m->dZ1, g.sai = 1, g.sap = m->hid, g.sbp = m->in, g.sbj = 1;
    g.M = m->hid, g.N = m->in, g.K = n, g.ldc = m->in, g.C = m->G + m->oW1;
    g.part = m->gpart, g.part_cap = m->gpart_cap;
    OBJ_TRY(m, ccd::launch_gemm(g, m->Xn, false, ccd::GEMM_STORE, s));
    OBJ_TRY(m, ccd::launch_colsum(m->dZ1, n, m->hid, m->G + m->ob1, s));
    return CCD_OK;
}

int mlp_sync_bf16(ccd_mlp* m, hipStream_t s) {
    if (m->p16_valid) return CCD_OK;
    OBJ_TRY(m, ccd::launch_to_bf16(m->P, m->P16, m->np, s));
    m->p16_valid = true;
    return CCD_OK;
}

// Packed weights of the matrix-core path (allocated on first use).
int mlp_sync_mfma(ccd_mlp* m, hipStream_t s) {
    if (m->hid != ccd::MLP_MFMA_HIDDEN || ((m->in + 31) & ~31u) > ccd::MLP_MFMA_MAX_KIN ||
        m->out > ccd::MLP_MFMA_HIDDEN)
        return obj_err(m->err, CCD_ERR_UNSUPPORTED, "matrix-core inference needs 256 hidden units, in <= 288, out <= 256");
    if (m->pk_valid) return CCD_OK;
    if (!m->pk) {
        m->kin = (m->in + 31) & ~31u;
        m->nout = (m->out + 15) & ~15u;
        const size_t H = ccd::MLP_MFMA_HIDDEN;
        OBJ_TRY(m, dalloc(&m->pk, H * m->kin + H * H + (size_t)m->nout * H));
        OBJ_TRY(m, dalloc(&m->pkb, 2 * H + m->nout));
    }
    const size_t H = ccd::MLP_MFMA_HIDDEN;
    uint16_t* W1 = m->pk;
    uint16_t* W2 = W1 + H * m->kin;
    uint16_t* W3 = W2 + H * H;
    OBJ_TRY(m, ccd::launch_mlp_pack(m->P, m->in, m->out, m->kin, m->nout, W1, W2, W3, m->pkb, m->pkb + H,
                                    m->pkb + 2 * H, s));
    m->pk_valid = true;
    return CCD_OK;
}

// Y (device, raw) = MLP(X (device, raw)) for any row count.  mode 0: FP32;
// 1: bfloat16 weights, FP32 arithmetic (chunked GEMMs through the
// workspace); 2: the matrix-core path, one launch (bfloat16 weights and
// activations, FP32 accumulation).
int mlp_forward_dev(ccd_mlp* m, uint64_t rows, const float* X, float* Y, int mode, hipStream_t s) {
    if (mode == 2) {
        if (int rc = mlp_sync_mfma(m, s)) return rc;
        if (rows > 0xffffffffull) return obj_err(m->err, CCD_ERR_INVALID, "too many rows");
        const size_t H = ccd::MLP_MFMA_HIDDEN;
        ccd::MlpInferArgs g{};
        g.X = X, g.Y = Y, g.rows = (uint32_t)rows, g.in = m->in, g.out = m->out, g.kin = m->kin, g.nout = m->nout;
        g.W1 = m->pk, g.W2 = m->pk + H * m->kin, g.W3 = m->pk + H * m->kin + H * H;
        g.b1 = m->pkb, g.b2 = m->pkb + H, g.b3 = m->pkb + 2 * H;
        if (m->has_norm) g.in_sub = m->in_sub, g.in_mul = m->in_mul, g.out_sub = m->out_sub, g.out_mul = m->out_mul;
        OBJ_TRY(m, ccd::launch_mlp_infer(g, s));
        return CCD_OK;
    }
    const bool b16 = mode == 1;
    const uint32_t chunk = 16384;
    if (int rc = mlp_workspace(m, chunk)) return rc;
    if (b16)
        if (int rc = mlp_sync_bf16(m, s)) return rc;
    for (uint64_t r0 = 0; r0 < rows; r0 += chunk) {
        const uint32_t n = (uint32_t)std::min<uint64_t>(chunk, rows - r0);
        OBJ_TRY(m, ccd::launch_affine_rows(X + r0 * m->in, nullptr, n, m->in, m->has_norm ? m->in_sub : nullptr,
                                           m->in_mul, m->Xn, s));
        if (int rc = mlp_forward_ws(m, n, b16, s)) return rc;
        OBJ_TRY(m, ccd::launch_affine_rows(m->Yn, nullptr, n, m->out, m->has_norm ? m->out_sub : nullptr, m->out_mul,
                                           Y + r0 * m->out, s));
    }
    return CCD_OK;
}

}  // namespace

extern "C" {

int32_t ccd_mlp_create(int32_t device, uint32_t in, uint32_t hidden, uint32_t out, uint64_t seed, ccd_mlp** mo) {
    if (!mo) return CCD_ERR_INVALID;
    *mo = nullptr;
    if (in == 0 || hidden == 0 || out == 0 || in > 4096 || hidden > 4096 || out > 4096) return CCD_ERR_INVALID;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return CCD_ERR_HIP;
    if (device < 0 || device >= count) return CCD_ERR_INVALID;
    ccd_mlp* m = new (std::nothrow) ccd_mlp();
    if (!m) return CCD_ERR_ALLOC;
    m->device = device;
    m->in = in, m->hid = hidden, m->out = out;
    m->oW1 = 0;
    m->ob1 = m->oW1 + (size_t)hidden * in;
    m->oW2 = m->ob1 + hidden;
    m->ob2 = m->oW2 + (size_t)hidden * hidden;
    m->oW3 = m->ob2 + hidden;
    m->ob3 = m->oW3 + (size_t)out * hidden;
    m->np = m->ob3 + out;
    m->gpart_cap = 16ull * std::max<uint64_t>({(uint64_t)hidden * hidden, (uint64_t)hidden * in,
                                                (uint64_t)out * hidden});
    // Glorot-uniform weights, zero biases, seeded
    std::vector<float> P(m->np, 0.0f);
    uint64_t st = seed ^ 0x6d6c7021ull;
    auto fill = [&](size_t off, uint32_t fo, uint32_t fi) {
        const double a = std::sqrt(6.0 / (double)(fi + fo));
        for (size_t q = 0; q < (size_t)fo * fi; ++q) P[off + q] = (float)((2.0 * uniform01(st) - 1.0) * a);
    };
    fill(m->oW1, hidden, in);
    fill(m->oW2, hidden, hidden);
    fill(m->oW3, out, hidden);
    bool ok = hipSetDevice(device) == hipSuccess && hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking) == hipSuccess &&
              dalloc(&m->P, m->np) == hipSuccess && dalloc(&m->P16, m->np) == hipSuccess &&
              dalloc(&m->G, m->np) == hipSuccess && dalloc(&m->mom, m->np) == hipSuccess &&
              dalloc(&m->vel, m->np) == hipSuccess && dalloc(&m->in_sub, in) == hipSuccess &&
              dalloc(&m->in_mul, in) == hipSuccess && dalloc(&m->out_sub, out) == hipSuccess &&
              dalloc(&m->out_mul, out) == hipSuccess && dalloc(&m->tgt_sub, out) == hipSuccess &&
              dalloc(&m->tgt_mul, out) == hipSuccess && dalloc(&m->part, ccd::MSE_PARTS) == hipSuccess &&
              dalloc(&m->gpart, m->gpart_cap) == hipSuccess &&
              ccd::copy_sync(m->stream, m->P, P.data(), sizeof(float) * m->np, hipMemcpyHostToDevice) == hipSuccess &&
              ccd::launch_fill(m->mom, 0.0f, m->np, m->stream) == hipSuccess &&
              ccd::launch_fill(m->vel, 0.0f, m->np, m->stream) == hipSuccess &&
              hipStreamSynchronize(m->stream) == hipSuccess;
    if (!ok) {
        ccd_mlp_destroy(m);
        return CCD_ERR_HIP;
    }
    *mo = m;
    return CCD_OK;
}

int32_t ccd_mlp_destroy(ccd_mlp* m) {
    if (!m) return CCD_OK;
    (void)hipSetDevice(m->device);
    if (m->stream) (void)hipStreamSynchronize(m->stream);
    void* bufs[] = {m->P, m->P16, m->G, m->mom, m->vel, m->in_sub, m->in_mul, m->out_sub, m->out_mul, m->tgt_sub,
                    m->tgt_mul, m->gpart, m->Xn, m->Z1,
                    m->A1, m->Z2, m->A2, m->Yn, m->Tn, m->dY, m->dZ2, m->dZ1, m->part, m->loss, m->X, m->T, m->perm,
                    m->pk, m->pkb};
    for (void* b : bufs) dfree(b);
    if (m->stream) (void)hipStreamDestroy(m->stream);
    delete m;
    return CCD_OK;
}

const char* ccd_mlp_last_error(ccd_mlp* m) { return m ? m->err.c_str() : "null mlp"; }

uint64_t ccd_mlp_param_count(ccd_mlp* m) { return m ? m->np : 0; }

int32_t ccd_mlp_get_params(ccd_mlp* m, float* P) {
    if (!m || !P) return CCD_ERR_INVALID;
    OBJ_TRY(m, hipSetDevice(m->device));
    OBJ_TRY(m, hipStreamSynchronize(m->stream));
    OBJ_TRY(m, hipMemcpy(P, m->P, sizeof(float) * m->np, hipMemcpyDeviceToHost));
    return CCD_OK;
}

int32_t ccd_mlp_set_params(ccd_mlp* m, const float* P) {
    if (!m || !P) return CCD_ERR_INVALID;
    OBJ_TRY(m, hipSetDevice(m->device));
    OBJ_TRY(m, hipStreamSynchronize(m->stream));
    OBJ_TRY(m, ccd::copy_sync(m->stream, m->P, P, sizeof(float) * m->np, hipMemcpyHostToDevice));
    m->p16_valid = false;
    m->pk_valid = false;
    return CCD_OK;
}

int32_t ccd_mlp_set_norm(ccd_mlp* m, const float* in_mean, const float* in_std, const float* out_mean,
                         const float* out_std) {
    if (!m || !in_mean || !in_std || !out_mean || !out_std) return CCD_ERR_INVALID;
    // x_n = (x - mean) / std;  y = y_n * std + mean = (y_n - (-mean / std)) * std
    std::vector<float> a(m->in), b(m->in), c(m->out), d(m->out), e(m->out), g(m->out);
    for (uint32_t j = 0; j < m->in; ++j) {
        const float sd = std::max(in_std[j], 1e-12f);
        if (!std::isfinite(in_mean[j]) || !std::isfinite(sd)) return obj_err(m->err, CCD_ERR_INVALID, "non-finite norm");
        a[j] = in_mean[j];
        b[j] = 1.0f / sd;
    }
    for (uint32_t j = 0; j < m->out; ++j) {
        const float sd = std::max(out_std[j], 1e-12f);
        if (!std::isfinite(out_mean[j]) || !std::isfinite(sd)) return obj_err(m->err, CCD_ERR_INVALID, "non-finite norm");
        c[j] = -out_mean[j] / sd;
        d[j] = sd;
        e[j] = out_mean[j];
        g[j] = 1.0f / sd;
    }
    OBJ_TRY(m, hipSetDevice(m->device));
    OBJ_TRY(m, hipStreamSynchronize(m->stream));
    OBJ_TRY(m, ccd::copy_sync(m->stream, m->in_sub, a.data(), 4 * m->in, hipMemcpyHostToDevice));
    OBJ_TRY(m, ccd::copy_sync(m->stream, m->in_mul, b.data(), 4 * m->in, hipMemcpyHostToDevice));
    OBJ_TRY(m, ccd::copy_sync(m->stream, m->out_sub, c.data(), 4 * m->out, hipMemcpyHostToDevice));
    OBJ_TRY(m, ccd::copy_sync(m->stream, m->out_mul, d.data(), 4 * m->out, hipMemcpyHostToDevice));
    OBJ_TRY(m, ccd::copy_sync(m->stream, m->tgt_sub, e.data(), 4 * m->out, hipMemcpyHostToDevice));
    OBJ_TRY(m, ccd::copy_sync(m->stream, m->tgt_mul, g.data(), 4 * m->out, hipMemcpyHostToDevice));
    m->has_norm = true;
    return CCD_OK;
}

int32_t ccd_mlp_forward(ccd_mlp* m, uint64_t rows, const float* X, float* Y, int32_t bf16_weights) {
    if (!m || (rows && (!X || !Y))) return CCD_ERR_INVALID;
    if (rows == 0) return CCD_OK;
    OBJ_TRY(m, hipSetDevice(m->device));
    float *dX = nullptr, *dYo = nullptr;
    OBJ_TRY(m, dalloc(&dX, rows * m->in));
    if (dalloc(&dYo, rows * m->out) != hipSuccess) {
        dfree(dX);
        return obj_err(m->err, CCD_ERR_ALLOC, "output buffer");
    }
    int rc = CCD_OK;
    if (ccd::copy_sync(m->stream, dX, X, 4 * rows * m->in, hipMemcpyHostToDevice) != hipSuccess) rc = CCD_ERR_HIP;
    if (bf16_weights < 0 || bf16_weights > 2) rc = obj_err(m->err, CCD_ERR_INVALID, "bf16_weights must be 0, 1 or 2");
    if (!rc) rc = mlp_forward_dev(m, rows, dX, dYo, bf16_weights, m->stream);
    if (!rc && hipStreamSynchronize(m->stream) != hipSuccess) rc = CCD_ERR_HIP;
    if (!rc && hipMemcpy(Y, dYo, 4 * rows * m->out, hipMemcpyDeviceToHost) != hipSuccess) rc = CCD_ERR_HIP;
    dfree(dX);
    dfree(dYo);
    if (rc == CCD_ERR_HIP && m->err.empty()) m->err = "copy or launch failed";
    return rc;
}

int32_t ccd_mlp_loss_grad(ccd_mlp* m, uint64_t rows, const float* Xn, const float* Tn, double* loss, float* grad) {
    if (!m || !Xn || !Tn || !loss || !grad || rows == 0 || rows > (1u << 20)) return CCD_ERR_INVALID;
    OBJ_TRY(m, hipSetDevice(m->device));
    if (int rc = mlp_workspace(m, rows)) return rc;
    if (!m->loss) OBJ_TRY(m, dalloc(&m->loss, 1));
    OBJ_TRY(m, ccd::copy_sync(m->stream, m->Xn, Xn, 4 * rows * m->in, hipMemcpyHostToDevice));
    OBJ_TRY(m, ccd::copy_sync(m->stream, m->Tn, Tn, 4 * rows * m->out, hipMemcpyHostToDevice));
    if (int rc = mlp_forward_ws(m, (uint32_t)rows, false, m->stream)) return rc;
    if (int rc = mlp_backward_ws(m, (uint32_t)rows, m->loss, m->stream)) return rc;
    OBJ_TRY(m, hipStreamSynchronize(m->stream));
    OBJ_TRY(m, hipMemcpy(loss, m->loss, sizeof(double), hipMemcpyDeviceToHost));
    OBJ_TRY(m, hipMemcpy(grad, m->G, 4 * m->np, hipMemcpyDeviceToHost));
    return CCD_OK;
}

int32_t ccd_mlp_train(ccd_mlp* m, uint64_t rows, const float* X, const float* T, uint32_t epochs, uint32_t batch,
                      float lr, uint64_t seed, double* epoch_loss) {
    if (!m || !X || !T || rows == 0 || batch == 0 || batch > 65536 || rows > 0xffffffffull || !(lr > 0.0f))
        return CCD_ERR_INVALID;
    OBJ_TRY(m, hipSetDevice(m->device));
    hipStream_t s = m->stream;
    if (int rc = mlp_workspace(m, batch)) return rc;
    OBJ_TRY(m, ensure(&m->X, &m->cap_x, rows * m->in));
    OBJ_TRY(m, ensure(&m->T, &m->cap_t, rows * m->out));
    OBJ_TRY(m, ensure(&m->perm, &m->cap_perm, rows));
    const uint64_t nb = (rows + batch - 1) / batch;
    OBJ_TRY(m, ensure(&m->loss, &m->cap_loss, nb));
    OBJ_TRY(m, ccd::copy_sync(m->stream, m->X, X, 4 * rows * m->in, hipMemcpyHostToDevice));
    OBJ_TRY(m, ccd::copy_sync(m->stream, m->T, T, 4 * rows * m->out, hipMemcpyHostToDevice));
    std::vector<uint32_t> perm(rows);
    std::vector<double> losses(nb);
    uint64_t st = seed ^ 0x747261696eull;
    const float b1 = 0.9f, b2 = 0.999f;
    double last = NAN;
    for (uint32_t ep = 0; ep < epochs; ++ep) {
        std::iota(perm.begin(), perm.end(), 0u);
        for (uint64_t i = rows - 1; i > 0; --i) std::swap(perm[i], perm[splitmix(st) % (i + 1)]);
        OBJ_TRY(m, hipStreamSynchronize(s));
        OBJ_TRY(m, ccd::copy_sync(m->stream, m->perm, perm.data(), 4 * rows, hipMemcpyHostToDevice));
        for (uint64_t q = 0; q < nb; ++q) {
            const uint32_t n = (uint32_t)std::min<uint64_t>(batch, rows - q * batch);
            const uint32_t* pr = m->perm + q * batch;
            OBJ_TRY(m, ccd::launch_affine_rows(m->X, pr, n, m->in, m->has_norm ? m->in_sub : nullptr, m->in_mul, m->Xn, s));
            OBJ_TRY(m, ccd::launch_affine_rows(m->T, pr, n, m->out, m->has_norm ? m->tgt_sub : nullptr, m->tgt_mul,
                                               m->Tn, s));
            if (int rc = mlp_forward_ws(m, n, false, s)) return rc;
            if (int rc = mlp_backward_ws(m, n, m->loss + q, s)) return rc;
            ++m->step;
            ccd::AdamArgs ad{m->P, m->G, m->mom, m->vel, m->np, lr, b1, b2, 1e-8f,
                             1.0f - std::pow(b1, (float)m->step), 1.0f - std::pow(b2, (float)m->step)};
            OBJ_TRY(m, ccd::launch_adam(ad, s));
        }
        OBJ_TRY(m, hipStreamSynchronize(s));
        OBJ_TRY(m, hipMemcpy(losses.data(), m->loss, sizeof(double) * nb, hipMemcpyDeviceToHost));
        last = 0.0;
        for (double l : losses) last += l;
        last /= (double)nb;
    }
    m->p16_valid = false;
    m->pk_valid = false;
    if (epoch_loss) *epoch_loss = last;
    return CCD_OK;
}

}  // extern "C"

// ------------------------------------------------- allocation-free path ---
// Scratch of ccd_fold_rows_noalloc: one call's staged batch -- row pointer,
// long-row list, indices, values, the per-rating residual buffer of the long
// rows, start factors and result factors -- carved from one file-scope
// device array.  Every GPU the library runs on holds its own copy (about
// 28 MB) until the process ends.  Calls on one GPU share it: calls that
// overlap on one GPU race.  Nothing in it needs to start at zero (the sweep
// kernel seeds the residual buffer from the values, and W from W0 or 0).
namespace {
constexpr uint64_t NA_MAX_ROWS = 4096;
constexpr uint64_t NA_MAX_NNZ = 1ull << 21;
constexpr size_t na_round(size_t b) { return (b + 255) & ~(size_t)255; }
constexpr size_t NA_OFF_PTR = 0;
constexpr size_t NA_OFF_LONG = NA_OFF_PTR + na_round(4 * (NA_MAX_ROWS + 1));
constexpr size_t NA_OFF_IDX = NA_OFF_LONG + na_round(4 * NA_MAX_ROWS);
constexpr size_t NA_OFF_VAL = NA_OFF_IDX + na_round(4 * NA_MAX_NNZ);
constexpr size_t NA_OFF_RBUF = NA_OFF_VAL + na_round(4 * NA_MAX_NNZ);
constexpr size_t NA_OFF_W0 = NA_OFF_RBUF + na_round(4 * NA_MAX_NNZ);
constexpr size_t NA_OFF_W = NA_OFF_W0 + na_round(4 * NA_MAX_ROWS * ccd::FOLD_KMAX);
constexpr size_t NA_BYTES = NA_OFF_W + na_round(4 * NA_MAX_ROWS * ccd::FOLD_KMAX);
}  // namespace

static __device__ __align__(256) unsigned char scratch[NA_BYTES];

namespace {
// Device address of `scratch` on each device, looked up once per device
// with that device current (the array's host-side name is no device
// address).
std::mutex g_scratch_mu;
std::vector<unsigned char*> g_scratch;

hipError_t scratch_on(int device, unsigned char** out) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if ((size_t)device >= g_scratch.size()) g_scratch.resize((size_t)device + 1, nullptr);
    if (!g_scratch[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch[device] = static_cast<unsigned char*>(p);
    }
    *out = g_scratch[device];
    return hipSuccess;
}
}  // namespace

// -------------------------------------------------------------- fold-in ---
struct ccd_fold {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    uint32_t k = 0;
    int64_t n = 0;
    float lambda = 0.0f;
    float *H = nullptr, *HT = nullptr;
    uint32_t *ptr = nullptr, *idx = nullptr, *long_rows = nullptr;
    uint32_t nlong = 0;
    float *val = nullptr, *rbuf = nullptr, *W0 = nullptr, *W = nullptr, *F = nullptr;
    size_t cap_ptr = 0, cap_idx = 0, cap_val = 0, cap_rbuf = 0, cap_w0 = 0, cap_w = 0, cap_f = 0, cap_long = 0;
    hipEvent_t ea = nullptr, eb = nullptr;
    unsigned char* scratch = nullptr;  // this device's scratch (ccd_fold_rows_noalloc)
};

namespace {

int fold_check(ccd_fold* f, int64_t rows, const uint32_t* ptr, const uint32_t* idx, const float* val) {
    if (rows < 0 || rows > 0x7fffffff || !ptr || (rows && !idx && ptr[rows] > 0) || (rows && !val && ptr[rows] > 0))
        return obj_err(f->err, CCD_ERR_INVALID, "bad rows / null arrays");
    if (ptr[0] != 0) return obj_err(f->err, CCD_ERR_INVALID, "ptr[0] must be 0");
    for (int64_t i = 0; i < rows; ++i)
        if (ptr[i + 1] < ptr[i]) return obj_err(f->err, CCD_ERR_INVALID, "ptr not monotone");
    const uint64_t nnz = ptr[rows];
    for (uint64_t e = 0; e < nnz; ++e)
        if ((int64_t)idx[e] >= f->n) return obj_err(f->err, CCD_ERR_INVALID, "column index out of range");
    return CCD_OK;
}

int fold_upload(ccd_fold* f, int64_t rows, const uint32_t* ptr, const uint32_t* idx, const float* val,
                ccd::FoldArgs* a) {
    if (int rc = fold_check(f, rows, ptr, idx, val)) return rc;
    const uint64_t nnz = ptr[rows];
    std::vector<uint32_t> lr;  // rows handled by a block each
    for (int64_t i = 0; i < rows; ++i)
        if (ptr[i + 1] - ptr[i] > ccd::FOLD_SHORT) lr.push_back((uint32_t)i);
    OBJ_TRY(f, hipSetDevice(f->device));
    OBJ_TRY(f, ensure(&f->ptr, &f->cap_ptr, (size_t)rows + 1));
    OBJ_TRY(f, ensure(&f->idx, &f->cap_idx, nnz));
    OBJ_TRY(f, ensure(&f->val, &f->cap_val, nnz));
    OBJ_TRY(f, ensure(&f->rbuf, &f->cap_rbuf, nnz));
    OBJ_TRY(f, ensure(&f->long_rows, &f->cap_long, lr.size()));
    OBJ_TRY(f, ccd::copy_sync(f->stream, f->ptr, ptr, 4 * ((size_t)rows + 1), hipMemcpyHostToDevice));
    if (!lr.empty()) OBJ_TRY(f, ccd::copy_sync(f->stream, f->long_rows, lr.data(), 4 * lr.size(), hipMemcpyHostToDevice));
    f->nlong = (uint32_t)lr.size();
    if (nnz) {
        OBJ_TRY(f, ccd::copy_sync(f->stream, f->idx, idx, 4 * nnz, hipMemcpyHostToDevice));
        OBJ_TRY(f, ccd::copy_sync(f->stream, f->val, val, 4 * nnz, hipMemcpyHostToDevice));
    }
    *a = ccd::FoldArgs{f->ptr, f->idx, f->val, f->H, f->HT, (uint32_t)rows, f->k, (uint32_t)f->n, 0, f->lambda};
    return CCD_OK;
}

}  // namespace

extern "C" {

int32_t ccd_fold_create(int32_t device, uint32_t k, int64_t n, const float* H, float lambda, ccd_fold** fo) {
    if (!fo) return CCD_ERR_INVALID;
    *fo = nullptr;
    if (k == 0 || k > (uint32_t)ccd::FOLD_KMAX || n <= 0 || n > 0x7fffffff || !H || !std::isfinite(lambda) ||
        lambda < 0.0f)
        return CCD_ERR_INVALID;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return CCD_ERR_HIP;
    if (device < 0 || device >= count) return CCD_ERR_INVALID;
    ccd_fold* f = new (std::nothrow) ccd_fold();
    if (!f) return CCD_ERR_ALLOC;
    f->device = device, f->k = k, f->n = n, f->lambda = lambda;
    const size_t kn = (size_t)k * n;
    bool ok = hipSetDevice(device) == hipSuccess && hipStreamCreateWithFlags(&f->stream, hipStreamNonBlocking) == hipSuccess &&
              hipEventCreate(&f->ea) == hipSuccess && hipEventCreate(&f->eb) == hipSuccess &&
              dalloc(&f->H, kn) == hipSuccess && dalloc(&f->HT, kn) == hipSuccess &&
              ccd::copy_sync(f->stream, f->H, H, 4 * kn, hipMemcpyHostToDevice) == hipSuccess &&
              ccd::launch_transpose(f->H, f->HT, k, (uint32_t)n, f->stream) == hipSuccess &&
              hipStreamSynchronize(f->stream) == hipSuccess && scratch_on(device, &f->scratch) == hipSuccess;
    if (!ok) {
        ccd_fold_destroy(f);
        return CCD_ERR_HIP;
    }
    *fo = f;
    return CCD_OK;
}

int32_t ccd_fold_destroy(ccd_fold* f) {
    if (!f) return CCD_OK;
    (void)hipSetDevice(f->device);
    if (f->stream) (void)hipStreamSynchronize(f->stream);
    void* bufs[] = {f->H, f->HT, f->ptr, f->idx, f->val, f->rbuf, f->W0, f->W, f->F, f->long_rows};
    for (void* b : bufs) dfree(b);
    if (f->ea) (void)hipEventDestroy(f->ea);
    if (f->eb) (void)hipEventDestroy(f->eb);
    if (f->stream) (void)hipStreamDestroy(f->stream);
    delete f;
    return CCD_OK;
}

const char* ccd_fold_last_error(ccd_fold* f) { return f ? f->err.c_str() : "null fold"; }

int32_t ccd_fold_features(ccd_fold* f, int64_t rows, const uint32_t* ptr, const uint32_t* idx, const float* val,
                          float* F) {
    if (!f || !F) return CCD_ERR_INVALID;
    ccd::FoldArgs a{};
    if (int rc = fold_upload(f, rows, ptr, idx, val, &a)) return rc;
    const size_t nf = (size_t)rows * (2 * f->k + 2);
    OBJ_TRY(f, ensure(&f->F, &f->cap_f, nf));
    OBJ_TRY(f, ccd::launch_fold_features(a, f->long_rows, f->nlong, f->F, f->stream));
    OBJ_TRY(f, hipStreamSynchronize(f->stream));
    if (nf) OBJ_TRY(f, hipMemcpy(F, f->F, 4 * nf, hipMemcpyDeviceToHost));
    return CCD_OK;
}

int32_t ccd_fold_rows(ccd_fold* f, int64_t rows, const uint32_t* ptr, const uint32_t* idx, const float* val,
                      const float* W0, int32_t sweeps, float* W, double* seconds) {
    if (!f || !W || sweeps < 0) return CCD_ERR_INVALID;
    ccd::FoldArgs a{};
    if (int rc = fold_upload(f, rows, ptr, idx, val, &a)) return rc;
    a.sweeps = sweeps;
    const size_t nw = (size_t)rows * f->k;
    OBJ_TRY(f, ensure(&f->W, &f->cap_w, nw));
    if (W0) {
        OBJ_TRY(f, ensure(&f->W0, &f->cap_w0, nw));
        if (nw) OBJ_TRY(f, ccd::copy_sync(f->stream, f->W0, W0, 4 * nw, hipMemcpyHostToDevice));
    }
    OBJ_TRY(f, hipEventRecord(f->ea, f->stream));
    OBJ_TRY(f, ccd::launch_fold_sweeps(a, f->long_rows, f->nlong, W0 ? f->W0 : nullptr, f->W, f->rbuf, f->stream));
    OBJ_TRY(f, hipEventRecord(f->eb, f->stream));
    OBJ_TRY(f, hipEventSynchronize(f->eb));
    float ms = 0.0f;
    OBJ_TRY(f, hipEventElapsedTime(&ms, f->ea, f->eb));
    if (seconds) *seconds = ms * 1e-3;
    if (nw) OBJ_TRY(f, hipMemcpy(W, f->W, 4 * nw, hipMemcpyDeviceToHost));
    return CCD_OK;
}

int32_t ccd_fold_rows_noalloc(ccd_fold* f, int64_t rows, const uint32_t* ptr, const uint32_t* idx, const float* val,
                              const float* W0, int32_t sweeps, float* W, double* seconds) {
    if (!f || !W || sweeps < 0) return CCD_ERR_INVALID;
    if (!f->scratch) return obj_err(f->err, CCD_ERR_STATE, "no scratch on this device");
    if (int rc = fold_check(f, rows, ptr, idx, val)) return rc;
    const uint64_t nnz = ptr[rows];
    if ((uint64_t)rows > NA_MAX_ROWS || nnz > NA_MAX_NNZ)
        return obj_err(f->err, CCD_ERR_INVALID,
                       "ccd_fold_rows_noalloc: at most 4096 rows and 2097152 ratings per call (got " +
                           std::to_string(rows) + " rows, " + std::to_string(nnz) + " ratings)");
    // long rows (one block each), listed in a fixed-size host array
    uint32_t lr[NA_MAX_ROWS];
    uint32_t nlong = 0;
    for (int64_t i = 0; i < rows; ++i)
        if (ptr[i + 1] - ptr[i] > ccd::FOLD_SHORT) lr[nlong++] = (uint32_t)i;
    unsigned char* S = f->scratch;
    uint32_t* dptr = reinterpret_cast<uint32_t*>(S + NA_OFF_PTR);
    uint32_t* dlong = reinterpret_cast<uint32_t*>(S + NA_OFF_LONG);
    uint32_t* didx = reinterpret_cast<uint32_t*>(S + NA_OFF_IDX);
    float* dval = reinterpret_cast<float*>(S + NA_OFF_VAL);
    float* drbuf = reinterpret_cast<float*>(S + NA_OFF_RBUF);
    float* dW0 = reinterpret_cast<float*>(S + NA_OFF_W0);
    float* dW = reinterpret_cast<float*>(S + NA_OFF_W);
    const size_t nw = (size_t)rows * f->k;
    OBJ_TRY(f, hipSetDevice(f->device));
    OBJ_TRY(f, ccd::copy_sync(f->stream, dptr, ptr, 4 * ((size_t)rows + 1), hipMemcpyHostToDevice));
    if (nlong) OBJ_TRY(f, ccd::copy_sync(f->stream, dlong, lr, 4 * (size_t)nlong, hipMemcpyHostToDevice));
    if (nnz) {
        OBJ_TRY(f, ccd::copy_sync(f->stream, didx, idx, 4 * nnz, hipMemcpyHostToDevice));
        OBJ_TRY(f, ccd::copy_sync(f->stream, dval, val, 4 * nnz, hipMemcpyHostToDevice));
    }
    if (W0 && nw) OBJ_TRY(f, ccd::copy_sync(f->stream, dW0, W0, 4 * nw, hipMemcpyHostToDevice));
    const ccd::FoldArgs a{dptr, didx, dval, f->H, f->HT, (uint32_t)rows, f->k, (uint32_t)f->n, sweeps, f->lambda};
    OBJ_TRY(f, hipEventRecord(f->ea, f->stream));
    OBJ_TRY(f, ccd::launch_fold_sweeps(a, dlong, nlong, W0 ? dW0 : nullptr, dW, drbuf, f->stream));
    OBJ_TRY(f, hipEventRecord(f->eb, f->stream));
    OBJ_TRY(f, hipEventSynchronize(f->eb));
    float ms = 0.0f;
    OBJ_TRY(f, hipEventElapsedTime(&ms, f->ea, f->eb));
    if (seconds) *seconds = ms * 1e-3;
    if (nw) OBJ_TRY(f, hipMemcpy(W, dW, 4 * nw, hipMemcpyDeviceToHost));
    return CCD_OK;
}

int32_t ccd_fold_warm(ccd_fold* f, ccd_mlp* m, int64_t rows, const uint32_t* ptr, const uint32_t* idx,
                      const float* val, int32_t sweeps, int32_t bf16_weights, float* W, double* seconds) {
    if (!f || !m || !W || sweeps < 0) return CCD_ERR_INVALID;
    if (m->in != 2 * f->k + 2 || m->out != f->k)
        return obj_err(f->err, CCD_ERR_INVALID, "the MLP must map 2k + 2 features to k outputs");
    if (m->device != f->device) return obj_err(f->err, CCD_ERR_INVALID, "MLP on another device");
    ccd::FoldArgs a{};
    if (int rc = fold_upload(f, rows, ptr, idx, val, &a)) return rc;
    a.sweeps = sweeps;
    const size_t nw = (size_t)rows * f->k, nf = (size_t)rows * (2 * f->k + 2);
    OBJ_TRY(f, ensure(&f->W, &f->cap_w, nw));
    OBJ_TRY(f, ensure(&f->W0, &f->cap_w0, nw));
    OBJ_TRY(f, ensure(&f->F, &f->cap_f, nf));
    OBJ_TRY(f, hipStreamSynchronize(m->stream));
    OBJ_TRY(f, hipEventRecord(f->ea, f->stream));
    OBJ_TRY(f, ccd::launch_fold_features(a, f->long_rows, f->nlong, f->F, f->stream));
    if (bf16_weights < 0 || bf16_weights > 2) return obj_err(f->err, CCD_ERR_INVALID, "bf16_weights must be 0, 1 or 2");
    if (int rc = mlp_forward_dev(m, (uint64_t)rows, f->F, f->W0, bf16_weights, f->stream))
        return obj_err(f->err, rc, "MLP: " + m->err);
    OBJ_TRY(f, ccd::launch_fold_sweeps(a, f->long_rows, f->nlong, f->W0, f->W, f->rbuf, f->stream));
    OBJ_TRY(f, hipEventRecord(f->eb, f->stream));
    OBJ_TRY(f, hipEventSynchronize(f->eb));
    float ms = 0.0f;
    OBJ_TRY(f, hipEventElapsedTime(&ms, f->ea, f->eb));
    if (seconds) *seconds = ms * 1e-3;
    if (nw) OBJ_TRY(f, hipMemcpy(W, f->W, 4 * nw, hipMemcpyDeviceToHost));
    return CCD_OK;
}

}  // extern "C"
