// Host data layer: data-directory reader/writer and seeded generators.
// See include/cdmf_data.h for the interface and the reference functions
// each entry point stands in for.
#include "cdmf_data.h"

#include <omp.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <string>
#include <memory>
#include <unordered_set>
#include <vector>

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

// ---------------------------------------------------------------- PRNG --
// SplitMix64 for seeding, xoshiro256** for the streams (Blackman & Vigna).
inline uint64_t splitmix64(uint64_t& x) {
    uint64_t z = (x += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

struct Xoshiro {
    uint64_t s[4];
    explicit Xoshiro(uint64_t seed) {
        uint64_t x = seed;
        for (auto& v : s) v = splitmix64(x);
    }
    static inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    inline uint64_t next() {
        const uint64_t r = rotl(s[1] * 5, 7) * 9;
        const uint64_t t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return r;
    }
    inline double uniform() { return (double)(next() >> 11) * 0x1.0p-53; }  // [0,1)
    inline uint64_t below(uint64_t n) {  // unbiased in [0, n)
        uint64_t lim = (uint64_t)(-n) % n;
        for (;;) {
            uint64_t r = next();
            if (r >= lim) return r % n;
        }
    }
};

inline uint64_t stream_seed(uint64_t seed, uint64_t tag, uint64_t idx) {
    uint64_t x = seed ^ (tag * 0xd1b54a32d192ed03ull) ^ (idx * 0x9e3779b97f4a7c15ull);
    return splitmix64(x);
}

// ------------------------------------------------------------ ratings ---
struct RatingTable {
    std::vector<double> cdf;
    std::vector<float> value;
    explicit RatingTable(int kind) {
        std::vector<double> p;
        if (kind == CDMF_RATING_NETFLIX_1_5 || kind == CDMF_RATING_PLANTED_1_5) {
            p = {0.05, 0.10, 0.29, 0.34, 0.22};
            for (int r = 1; r <= 5; ++r) value.push_back((float)r);
        } else if (kind == CDMF_RATING_HALF_STAR) {
            // 0.5 .. 5.0: frequencies shaped like MovieLens-20M, mode at 4.0
            p = {0.011, 0.033, 0.012, 0.072, 0.044, 0.214, 0.110, 0.278, 0.077, 0.149};
            for (int r = 1; r <= 10; ++r) value.push_back(0.5f * (float)r);
        } else {
            p = {0.2, 0.2, 0.2, 0.2, 0.2};
            for (int r = 1; r <= 5; ++r) value.push_back((float)r);
        }
        double s = std::accumulate(p.begin(), p.end(), 0.0), c = 0.0;
        for (double q : p) { c += q / s; cdf.push_back(c); }
        cdf.back() = 1.0;
    }
    inline float draw(Xoshiro& g) const { return at(g.uniform()); }
    // the rating whose cumulative frequency band holds u in [0, 1)
    inline float at(double u) const {
        size_t i = 0;
        while (i + 1 < cdf.size() && u >= cdf[i]) ++i;
        return value[i];
    }
};

// Hidden model of CDMF_RATING_PLANTED_1_5: per row a_i, u_i (PLANT_K values),
// per column b_j, v_j; score = a_i + b_j + u_i.v_j + e, every term Gaussian,
// so score / sd is close to N(0, 1) and Phi(score / sd) close to uniform.
constexpr int PLANT_K = 8;
constexpr double PLANT_SA = 0.5, PLANT_SB = 0.5, PLANT_SU = 0.5, PLANT_SE = 0.5;

inline double gauss(Xoshiro& g) {  // Box-Muller
    const double u1 = std::max(g.uniform(), 1e-300), u2 = g.uniform();
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
}

struct Planted {
    std::vector<float> a, u, b, v;
    double sd = 1.0;
    Planted(int64_t m, int64_t n, uint64_t seed) : a(m), u((size_t)m * PLANT_K), b(n), v((size_t)n * PLANT_K) {
        for (int64_t i = 0; i < m; ++i) {
            Xoshiro g(stream_seed(seed, 0x7a11, (uint64_t)i));
            a[i] = (float)(PLANT_SA * gauss(g));
            for (int q = 0; q < PLANT_K; ++q) u[(size_t)i * PLANT_K + q] = (float)(PLANT_SU * gauss(g));
        }
        for (int64_t j = 0; j < n; ++j) {
            Xoshiro g(stream_seed(seed, 0x7b22, (uint64_t)j));
            b[j] = (float)(PLANT_SB * gauss(g));
            for (int q = 0; q < PLANT_K; ++q) v[(size_t)j * PLANT_K + q] = (float)(PLANT_SU * gauss(g));
        }
        sd = std::sqrt(PLANT_SA * PLANT_SA + PLANT_SB * PLANT_SB + PLANT_K * std::pow(PLANT_SU, 4) + PLANT_SE * PLANT_SE);
    }
    float rating(const RatingTable& rt, int64_t i, uint32_t j, Xoshiro& g) const {
        double sc = (double)a[i] + (double)b[j] + PLANT_SE * gauss(g);
        for (int q = 0; q < PLANT_K; ++q) sc += (double)u[(size_t)i * PLANT_K + q] * v[(size_t)j * PLANT_K + q];
        return rt.at(0.5 * std::erfc(-(sc / sd) / std::sqrt(2.0)));
    }
};

// ------------------------------------------------------- alias table ----
struct Alias {
    std::vector<double> prob;
    std::vector<uint32_t> alias;
    explicit Alias(const std::vector<double>& w) {
        size_t n = w.size();
        prob.assign(n, 0.0);
        alias.assign(n, 0);
        double s = std::accumulate(w.begin(), w.end(), 0.0);
        std::vector<double> q(n);
        std::vector<uint32_t> small, large;
        for (size_t i = 0; i < n; ++i) {
            q[i] = w[i] * (double)n / s;
            (q[i] < 1.0 ? small : large).push_back((uint32_t)i);
        }
        while (!small.empty() && !large.empty()) {
            uint32_t a = small.back(); small.pop_back();
            uint32_t b = large.back();
            prob[a] = q[a];
            alias[a] = b;
            q[b] -= 1.0 - q[a];
            if (q[b] < 1.0) { large.pop_back(); small.push_back(b); }
        }
        for (uint32_t i : large) prob[i] = 1.0;
        for (uint32_t i : small) prob[i] = 1.0;
    }
    inline uint32_t draw(Xoshiro& g) const {
        uint64_t i = g.below(prob.size());
        return g.uniform() < prob[i] ? (uint32_t)i : alias[i];
    }
};

// --------------------------------------------------- dataset helpers ----
template <class T>
T* xalloc(size_t count) {
    return static_cast<T*>(std::malloc(sizeof(T) * (count ? count : 1)));
}

void zero_ds(cdmf_dataset* d) { std::memset(d, 0, sizeof(*d)); }

// Build train CSR/CSC and test COO from per-row sorted entries (row-major
// order) and a per-entry hold-out flag.
int assemble(int64_t m, int64_t n, const std::vector<uint64_t>& row_off, const std::vector<uint32_t>& cols,
             const std::vector<float>& vals, const std::vector<uint8_t>& is_test, cdmf_dataset* out) {
    const uint64_t total = row_off[m];
    uint64_t ntest = 0;
    for (uint64_t e = 0; e < total; ++e) ntest += is_test[e];
    const uint64_t ntrain = total - ntest;
    if (ntrain > 0xffffffffull) return fail(CDMF_ERR_INVALID, "nnz exceeds the u32 index range");
    zero_ds(out);
    out->m = m;
    out->n = n;
    out->nnz = ntrain;
    out->nnz_test = ntest;
    out->csr_ptr = xalloc<uint32_t>(m + 1);
    out->csr_idx = xalloc<uint32_t>(ntrain);
    out->csr_val = xalloc<float>(ntrain);
    out->csc_ptr = xalloc<uint32_t>(n + 1);
    out->csc_idx = xalloc<uint32_t>(ntrain);
    out->csc_val = xalloc<float>(ntrain);
    out->test_row = xalloc<uint32_t>(ntest);
    out->test_col = xalloc<uint32_t>(ntest);
    out->test_val = xalloc<float>(ntest);
    if (!out->csr_ptr || !out->csr_idx || !out->csr_val || !out->csc_ptr || !out->csc_idx || !out->csc_val ||
        !out->test_row || !out->test_col || !out->test_val) {
        cdmf_dataset_free(out);
        return fail(CDMF_ERR_ALLOC, "out of host memory");
    }
    // per-row train counts, then prefix sums (parallel count, serial scan)
    std::vector<uint64_t> rcount(m, 0), tcount(m, 0);
#pragma omp parallel for schedule(dynamic, 1024)
    for (int64_t i = 0; i < m; ++i) {
        uint64_t c = 0;
        for (uint64_t e = row_off[i]; e < row_off[i + 1]; ++e) c += !is_test[e];
        rcount[i] = c;
        tcount[i] = (row_off[i + 1] - row_off[i]) - c;
    }
    std::vector<uint64_t> toff(m + 1, 0);
    uint64_t acc = 0, tacc = 0;
    uint64_t maxr = 0;
    for (int64_t i = 0; i < m; ++i) {
        out->csr_ptr[i] = (uint32_t)acc;
        toff[i] = tacc;
        acc += rcount[i];
        tacc += tcount[i];
        maxr = std::max(maxr, rcount[i]);
    }
    out->csr_ptr[m] = (uint32_t)acc;
    toff[m] = tacc;
    out->max_row_nnz = maxr;
#pragma omp parallel for schedule(dynamic, 1024)
    for (int64_t i = 0; i < m; ++i) {
        uint64_t p = out->csr_ptr[i], q = toff[i];
        for (uint64_t e = row_off[i]; e < row_off[i + 1]; ++e) {
            if (is_test[e]) {
                out->test_row[q] = (uint32_t)i;
                out->test_col[q] = cols[e];
                out->test_val[q] = vals[e];
                ++q;
            } else {
                out->csr_idx[p] = cols[e];
                out->csr_val[p] = vals[e];
                ++p;
            }
        }
    }
    // CSC by counting sort; scanning rows in order keeps rows ascending per column
    std::vector<uint64_t> ccount(n + 1, 0);
    for (uint64_t e = 0; e < ntrain; ++e) ccount[out->csr_idx[e] + 1]++;
    uint64_t maxc = 0;
    for (int64_t j = 0; j < n; ++j) {
        maxc = std::max(maxc, ccount[j + 1]);
        ccount[j + 1] += ccount[j];
    }
    out->max_col_nnz = maxc;
    for (int64_t j = 0; j <= n; ++j) out->csc_ptr[j] = (uint32_t)ccount[j];
    std::vector<uint64_t> cur(ccount.begin(), ccount.end() - 1);
    for (int64_t i = 0; i < m; ++i)
        for (uint64_t e = out->csr_ptr[i]; e < out->csr_ptr[i + 1]; ++e) {
            uint32_t j = out->csr_idx[e];
            uint64_t p = cur[j]++;
            out->csc_idx[p] = (uint32_t)i;
            out->csc_val[p] = out->csr_val[e];
        }
    return CDMF_OK;
}

// Selection sampling (Knuth, TAOCP Vol. 2, Algorithm S): exactly `k` of `N`
// positions, uniformly, in one sequential pass.
void select_holdout(uint64_t N, uint64_t k, uint64_t seed, std::vector<uint8_t>& flag) {
    flag.assign(N, 0);
    Xoshiro g(stream_seed(seed, 0x7e57, 0));
    uint64_t chosen = 0;
    for (uint64_t t = 0; t < N && chosen < k; ++t) {
        if ((double)(N - t) * g.uniform() < (double)(k - chosen)) {
            flag[t] = 1;
            ++chosen;
        }
    }
}

// Mean of the discrete power law P(d) ~ d^-a on [1, D].
double powerlaw_mean(double a, uint32_t D) {
    double s0 = 0.0, s1 = 0.0;
    for (uint32_t d = 1; d <= D; ++d) {
        double w = std::pow((double)d, -a);
        s0 += w;
        s1 += w * d;
    }
    return s1 / s0;
}

}  // namespace

extern "C" {

const char* cdmf_data_last_error(void) { return g_err.c_str(); }

void cdmf_dataset_free(cdmf_dataset* d) {
    if (!d) return;
    std::free(d->csr_ptr);
    std::free(d->csr_idx);
    std::free(d->csr_val);
    std::free(d->csc_ptr);
    std::free(d->csc_idx);
    std::free(d->csc_val);
    std::free(d->test_row);
    std::free(d->test_col);
    std::free(d->test_val);
    zero_ds(d);
}

void cdmf_initial_col(float* X, uint32_t k, uint32_t n) {
    // tools.cpp:349-357 draws glibc rand() after srand(0), i-major.
    srand(0);
    for (uint32_t i = 0; i < n; ++i)
        for (uint32_t j = 0; j < k; ++j)
            X[(size_t)j * n + i] = 0.1f * ((float)rand() / (float)RAND_MAX) + 0.001f;
}

void cdmf_init_uniform(float* X, uint64_t count, float scale, uint64_t seed) {
    Xoshiro g(stream_seed(seed, 0x1a17, 0));
    for (uint64_t e = 0; e < count; ++e) X[e] = scale * (float)g.uniform();
}

int cdmf_gen_uniform(int64_t m, int64_t n, uint64_t nnz_train, uint64_t nnz_test, int32_t rating_kind,
                     uint64_t seed, cdmf_dataset* out) {
    if (!out || m <= 0 || n <= 0) return fail(CDMF_ERR_INVALID, "bad shape");
    const uint64_t total = nnz_train + nnz_test;
    const uint64_t cells = (uint64_t)m * (uint64_t)n;
    if (total > cells) return fail(CDMF_ERR_INVALID, "more entries than matrix cells");
    if (rating_kind == CDMF_RATING_PLANTED_1_5)
        return fail(CDMF_ERR_INVALID, "planted ratings come from the power-law generator only");
    // Floyd's algorithm: `total` distinct cells out of `cells`.
    Xoshiro g(stream_seed(seed, 0x0f10, 0));
    std::unordered_set<uint64_t> seen;
    seen.reserve(total * 2);
    std::vector<uint64_t> pos;
    pos.reserve(total);
    for (uint64_t j = cells - total; j < cells; ++j) {
        uint64_t t = g.below(j + 1);
        if (seen.insert(t).second) pos.push_back(t);
        else { seen.insert(j); pos.push_back(j); }
    }
    std::sort(pos.begin(), pos.end());
    RatingTable rt(rating_kind);
    Xoshiro gr(stream_seed(seed, 0x4a7e, 0));
    std::vector<uint64_t> row_off(m + 1, 0);
    std::vector<uint32_t> cols(total);
    std::vector<float> vals(total);
    for (uint64_t e = 0; e < total; ++e) {
        row_off[pos[e] / n + 1]++;
        cols[e] = (uint32_t)(pos[e] % n);
        vals[e] = rt.draw(gr);
    }
    for (int64_t i = 0; i < m; ++i) row_off[i + 1] += row_off[i];
    std::vector<uint8_t> is_test;
    select_holdout(total, nnz_test, seed, is_test);
    return assemble(m, n, row_off, cols, vals, is_test, out);
}

int cdmf_gen_stress(char shape, uint64_t seed, cdmf_dataset* out) {
    if (!out) return fail(CDMF_ERR_INVALID, "null output");
    int64_t m, n, dense_row;
    uint32_t fixed_col, lo, hi;  // every sparse row: fixed_col and one column drawn from [lo, hi)
    if (shape == 'a') {
        m = 64000, n = 150, dense_row = 0, fixed_col = 0, lo = 1, hi = 150;
    } else if (shape == 'b') {
        m = n = 1024000, dense_row = m - 1, fixed_col = (uint32_t)(n - 1), lo = 0, hi = (uint32_t)(n - 1);
    } else if (shape == 'c') {
        m = n = 1024000, dense_row = -1, fixed_col = 0, lo = 1, hi = (uint32_t)n;
    } else {
        return fail(CDMF_ERR_INVALID, "stress shape must be 'a', 'b' or 'c'");
    }
    RatingTable rt(CDMF_RATING_UNIFORM_1_5);
    // training ratings: the drawn column of each sparse row, one stream per row
    std::vector<uint32_t> drawn(m, 0);
    std::vector<float> r_fixed(m), r_drawn(m);
#pragma omp parallel for schedule(static, 4096)
    for (int64_t i = 0; i < m; ++i) {
        Xoshiro g(stream_seed(seed, 0x5a, (uint64_t)i));
        drawn[i] = lo + (uint32_t)g.below(hi - lo);
        r_fixed[i] = rt.draw(g);
        r_drawn[i] = rt.draw(g);
    }
    const uint64_t ntrain = (dense_row >= 0 ? (uint64_t)n + 2ull * (uint64_t)(m - 1) : 2ull * (uint64_t)m);
    const uint64_t ntest = ntrain / 100;
    // test positions: sparse rows only, a column of [lo, hi) other than the
    // row's drawn one (so apart from the row's training ratings), distinct
    std::vector<std::vector<std::pair<uint32_t, float>>> test(m);
    {
        Xoshiro g(stream_seed(seed, 0x7e5, 0));
        std::unordered_set<uint64_t> seen;
        uint64_t got = 0;
        while (got < ntest) {
            const int64_t i = (int64_t)g.below((uint64_t)m);
            if (i == dense_row) continue;
            const uint32_t j = lo + (uint32_t)g.below(hi - lo);
            if (j == drawn[i] || j == fixed_col) continue;
            if (!seen.insert((uint64_t)i * (uint64_t)n + j).second) continue;
            test[i].push_back({j, rt.draw(g)});
            ++got;
        }
    }
    std::vector<uint64_t> row_off(m + 1, 0);
    for (int64_t i = 0; i < m; ++i)
        row_off[i + 1] = row_off[i] + (i == dense_row ? (uint64_t)n : 2u) + test[i].size();
    const uint64_t total = row_off[m];
    std::vector<uint32_t> cols(total);
    std::vector<float> vals(total);
    std::vector<uint8_t> is_test(total, 0);
    Xoshiro gd(stream_seed(seed, 0xde5e, 0));
    for (int64_t i = 0; i < m; ++i) {
        std::vector<std::pair<uint32_t, std::pair<float, uint8_t>>> row;
        if (i == dense_row) {
            for (int64_t j = 0; j < n; ++j) row.push_back({(uint32_t)j, {rt.draw(gd), 0}});
        } else {
            row.push_back({fixed_col, {r_fixed[i], 0}});
            row.push_back({drawn[i], {r_drawn[i], 0}});
        }
        for (const auto& t : test[i]) row.push_back({t.first, {t.second, 1}});
        std::sort(row.begin(), row.end());
        uint64_t e = row_off[i];
        for (const auto& q : row) {
            cols[e] = q.first;
            vals[e] = q.second.first;
            is_test[e] = q.second.second;
            ++e;
        }
    }
    return assemble(m, n, row_off, cols, vals, is_test, out);
}

int cdmf_gen_powerlaw(const cdmf_gen_spec* spec, cdmf_dataset* out) {
    if (!spec || !out || spec->m <= 0 || spec->n <= 0) return fail(CDMF_ERR_INVALID, "bad shape");
    const int64_t m = spec->m, n = spec->n;
    const uint64_t total = spec->nnz_train + spec->nnz_test;
    uint32_t D = spec->max_row_degree ? spec->max_row_degree : (uint32_t)n;
    D = (uint32_t)std::min<int64_t>(D, n);
    if (total > (uint64_t)m * (uint64_t)D) return fail(CDMF_ERR_INVALID, "nnz exceeds m * max_row_degree");
    if (spec->threads > 0) omp_set_num_threads(spec->threads);
    const uint64_t seed = spec->seed;
    const double target_mean = (double)total / (double)m;

    // 1. row-degree exponent
    double a = spec->row_alpha;
    if (a <= 0.0) {
        double lo = 0.0, hi = 6.0;  // mean is decreasing in a
        if (powerlaw_mean(lo, D) < target_mean) a = 0.0;
        else {
            for (int it = 0; it < 60; ++it) {
                double mid = 0.5 * (lo + hi);
                if (powerlaw_mean(mid, D) > target_mean) lo = mid; else hi = mid;
            }
            a = 0.5 * (lo + hi);
        }
    }
    // 2. degrees by inverse CDF, per-row streams
    std::vector<double> cdf(D);
    {
        double s = 0.0;
        for (uint32_t d = 1; d <= D; ++d) { s += std::pow((double)d, -a); cdf[d - 1] = s; }
        for (auto& c : cdf) c /= s;
        cdf.back() = 1.0;
    }
    std::vector<uint32_t> deg(m);
#pragma omp parallel for schedule(static, 4096)
    for (int64_t i = 0; i < m; ++i) {
        Xoshiro g(stream_seed(seed, 0xde9, (uint64_t)i));
        double u = g.uniform();
        deg[i] = (uint32_t)(std::lower_bound(cdf.begin(), cdf.end(), u) - cdf.begin()) + 1;
    }
    // 3. rescale to the exact total, then fix the remainder one unit at a time
    {
        double s = 0.0;
        for (auto d : deg) s += d;
        double f = (double)total / s;
        for (auto& d : deg) {
            double x = std::floor((double)d * f + 0.5);
            d = (uint32_t)std::min<double>(std::max<double>(x, 1.0), (double)D);
        }
        int64_t diff = (int64_t)total;
        for (auto d : deg) diff -= d;
        Xoshiro g(stream_seed(seed, 0xf1c5, 0));
        uint64_t guard = 0;
        while (diff != 0 && guard < 400000000ull) {
            uint64_t i = g.below((uint64_t)m);
            if (diff > 0 && deg[i] < D) { deg[i]++; diff--; }
            else if (diff < 0 && deg[i] > 1) { deg[i]--; diff++; }
            ++guard;
        }
        if (diff != 0) return fail(CDMF_ERR_INVALID, "could not reach the requested nnz");
    }
    // 4. column popularity: rank^-s over a seeded permutation of the columns
    std::vector<uint32_t> perm(n);
    std::iota(perm.begin(), perm.end(), 0u);
    {
        Xoshiro g(stream_seed(seed, 0x9e4, 0));
        for (int64_t j = n - 1; j > 0; --j) std::swap(perm[j], perm[g.below((uint64_t)j + 1)]);
    }
    std::vector<double> w(n);
    for (int64_t r = 0; r < n; ++r) w[perm[r]] = std::pow((double)(r + 1), -spec->col_exponent);
    Alias alias(w);
    RatingTable rt(spec->rating_kind);
    std::unique_ptr<Planted> plant;
    if (spec->rating_kind == CDMF_RATING_PLANTED_1_5) plant.reset(new Planted(m, n, seed));

    std::vector<uint64_t> row_off(m + 1, 0);
    for (int64_t i = 0; i < m; ++i) row_off[i + 1] = row_off[i] + deg[i];
    std::vector<uint32_t> cols(total);
    std::vector<float> vals(total);
    const uint32_t dense_cut = (uint32_t)std::max<int64_t>(64, n / 8);

#pragma omp parallel
    {
        std::vector<uint8_t> used(n, 0);
        std::vector<std::pair<double, uint32_t>> keys;
#pragma omp for schedule(dynamic, 256)
        for (int64_t i = 0; i < m; ++i) {
            Xoshiro g(stream_seed(seed, 0xc01, (uint64_t)i));
            const uint32_t d = deg[i];
            uint32_t* c = &cols[row_off[i]];
            if (d <= dense_cut) {
                // successive weighted draws, duplicates rejected
                uint32_t got = 0;
                while (got < d) {
                    uint32_t j = alias.draw(g);
                    if (!used[j]) { used[j] = 1; c[got++] = j; }
                }
                for (uint32_t q = 0; q < d; ++q) used[c[q]] = 0;
            } else {
                // Efraimidis-Spirakis: the d largest keys U^(1/w_j)
                keys.resize(n);
                for (int64_t j = 0; j < n; ++j) {
                    double u = g.uniform();
                    keys[j] = {std::log(u > 0.0 ? u : 1e-300) / w[j], (uint32_t)j};
                }
                std::nth_element(keys.begin(), keys.begin() + d, keys.end(),
                                 [](const std::pair<double, uint32_t>& x, const std::pair<double, uint32_t>& y) {
                                     return x.first > y.first;
                                 });
                for (uint32_t q = 0; q < d; ++q) c[q] = keys[q].second;
            }
            std::sort(c, c + d);
            Xoshiro gr(stream_seed(seed, 0x4a7e, (uint64_t)i));
            if (plant)
                for (uint32_t q = 0; q < d; ++q) vals[row_off[i] + q] = plant->rating(rt, i, c[q], gr);
            else
                for (uint32_t q = 0; q < d; ++q) vals[row_off[i] + q] = rt.draw(gr);
        }
    }
    std::vector<uint8_t> is_test;
    select_holdout(total, spec->nnz_test, seed, is_test);
    return assemble(m, n, row_off, cols, vals, is_test, out);
}

// ---------------------------------------------------------------- I/O ---

static bool read_exact(const std::string& fname, void* dst, size_t bytes, std::string& why) {
    FILE* f = std::fopen(fname.c_str(), "rb");
    if (!f) { why = "cannot open " + fname; return false; }
    size_t got = bytes ? std::fread(dst, 1, bytes, f) : 0;
    std::fclose(f);
    if (got != bytes) {
        char buf[256];
        std::snprintf(buf, sizeof buf, "FREAD read %zu bytes, needed %zu (%s)", got, bytes, fname.c_str());
        why = buf;
        return false;
    }
    return true;
}

// Pointer files may be shorter than count (scipy infers the shape from the
// largest index); missing entries are nnz.
static bool read_ptr(const std::string& fname, uint32_t* dst, uint64_t count, uint64_t nnz, uint64_t* maxdeg,
                     std::string& why) {
    FILE* f = std::fopen(fname.c_str(), "rb");
    if (!f) { why = "cannot open " + fname; return false; }
    size_t got = std::fread(dst, sizeof(uint32_t), count, f);
    std::fclose(f);
    for (uint64_t i = got; i < count; ++i) dst[i] = (uint32_t)nnz;
    uint64_t mx = 0;
    for (uint64_t i = 1; i < count; ++i) mx = std::max<uint64_t>(mx, (uint64_t)(dst[i] - dst[i - 1]));
    *maxdeg = mx;
    return true;
}

int cdmf_load_dir(const char* dir, cdmf_dataset* out) {
    if (!dir || !out) return fail(CDMF_ERR_INVALID, "null argument");
    zero_ds(out);
    std::string base(dir);
    FILE* fp = std::fopen((base + "/meta_modified_all").c_str(), "r");
    if (!fp) return fail(CDMF_ERR_IO, "Can't open meta input file.");
    long m = 0, n = 0;
    unsigned long nnz = 0, nnz_test = 0;
    char names[12][1024];
    bool ok = std::fscanf(fp, "%ld %ld %lu", &m, &n, &nnz) == 3;
    for (int q = 0; ok && q < 9; ++q) ok = std::fscanf(fp, "%1023s", names[q]) == 1;
    ok = ok && std::fscanf(fp, "%lu", &nnz_test) == 1;
    for (int q = 9; ok && q < 12; ++q) ok = std::fscanf(fp, "%1023s", names[q]) == 1;
    std::fclose(fp);
    if (!ok || m <= 0 || n <= 0) return fail(CDMF_ERR_IO, "malformed meta_modified_all");
    out->m = m;
    out->n = n;
    out->nnz = nnz;
    out->nnz_test = nnz_test;
    out->csr_ptr = xalloc<uint32_t>(m + 1);
    out->csr_idx = xalloc<uint32_t>(nnz);
    out->csr_val = xalloc<float>(nnz);
    out->csc_ptr = xalloc<uint32_t>(n + 1);
    out->csc_idx = xalloc<uint32_t>(nnz);
    out->csc_val = xalloc<float>(nnz);
    out->test_row = xalloc<uint32_t>(nnz_test);
    out->test_col = xalloc<uint32_t>(nnz_test);
    out->test_val = xalloc<float>(nnz_test);
    std::string why;
    auto p = [&](int q) { return base + "/" + names[q]; };
    ok = out->csr_ptr && out->csr_idx && out->csr_val && out->csc_ptr && out->csc_idx && out->csc_val &&
         out->test_row && out->test_col && out->test_val;
    if (!ok) { cdmf_dataset_free(out); return fail(CDMF_ERR_ALLOC, "out of host memory"); }
    ok = read_ptr(p(3), out->csr_ptr, m + 1, nnz, &out->max_row_nnz, why) &&
         read_exact(p(4), out->csr_idx, 4 * nnz, why) && read_exact(p(5), out->csr_val, 4 * nnz, why) &&
         read_ptr(p(6), out->csc_ptr, n + 1, nnz, &out->max_col_nnz, why) &&
         read_exact(p(7), out->csc_idx, 4 * nnz, why) && read_exact(p(8), out->csc_val, 4 * nnz, why) &&
         read_exact(p(9), out->test_val, 4 * nnz_test, why) && read_exact(p(10), out->test_row, 4 * nnz_test, why) &&
         read_exact(p(11), out->test_col, 4 * nnz_test, why);
    if (!ok) { cdmf_dataset_free(out); return fail(CDMF_ERR_IO, why); }
    // structural checks the reference leaves to luck
    for (long i = 0; i < m; ++i)
        if (out->csr_ptr[i + 1] < out->csr_ptr[i] || out->csr_ptr[i + 1] > nnz) {
            cdmf_dataset_free(out);
            return fail(CDMF_ERR_IO, "CSR pointer not monotone or beyond nnz");
        }
    for (long j = 0; j < n; ++j)
        if (out->csc_ptr[j + 1] < out->csc_ptr[j] || out->csc_ptr[j + 1] > nnz) {
            cdmf_dataset_free(out);
            return fail(CDMF_ERR_IO, "CSC pointer not monotone or beyond nnz");
        }
    for (unsigned long e = 0; e < nnz; ++e)
        if (out->csr_idx[e] >= (uint32_t)n || out->csc_idx[e] >= (uint32_t)m) {
            cdmf_dataset_free(out);
            return fail(CDMF_ERR_IO, "index out of range");
        }
    for (unsigned long e = 0; e < nnz_test; ++e)
        if (out->test_row[e] >= (uint32_t)m || out->test_col[e] >= (uint32_t)n) {
            cdmf_dataset_free(out);
            return fail(CDMF_ERR_IO, "test index out of range");
        }
    return CDMF_OK;
}

static bool write_bin(const std::string& fname, const void* src, size_t bytes) {
    FILE* f = std::fopen(fname.c_str(), "wb");
    if (!f) return false;
    bool ok = bytes == 0 || std::fwrite(src, 1, bytes, f) == bytes;
    return std::fclose(f) == 0 && ok;
}

int cdmf_write_dir(const char* dir, const cdmf_dataset* d) {
    if (!dir || !d) return fail(CDMF_ERR_INVALID, "null argument");
    std::string b(dir);
    const uint64_t nnz = d->nnz, nt = d->nnz_test;
    // COO triplets of the training set in CSR order
    std::vector<uint32_t> coo_row(nnz);
    for (int64_t i = 0; i < d->m; ++i)
        for (uint64_t e = d->csr_ptr[i]; e < d->csr_ptr[i + 1]; ++e) coo_row[e] = (uint32_t)i;
    bool ok = write_bin(b + "/R_train_coo.data.bin", d->csr_val, 4 * nnz) &&
              write_bin(b + "/R_train_coo.row.bin", coo_row.data(), 4 * nnz) &&
              write_bin(b + "/R_train_coo.col.bin", d->csr_idx, 4 * nnz) &&
              write_bin(b + "/R_train_csr.indptr.bin", d->csr_ptr, 4 * (d->m + 1)) &&
              write_bin(b + "/R_train_csr.indices.bin", d->csr_idx, 4 * nnz) &&
              write_bin(b + "/R_train_csr.data.bin", d->csr_val, 4 * nnz) &&
              write_bin(b + "/R_train_csc.indptr.bin", d->csc_ptr, 4 * (d->n + 1)) &&
              write_bin(b + "/R_train_csc.indices.bin", d->csc_idx, 4 * nnz) &&
              write_bin(b + "/R_train_csc.data.bin", d->csc_val, 4 * nnz) &&
              write_bin(b + "/R_test_coo.data.bin", d->test_val, 4 * nt) &&
              write_bin(b + "/R_test_coo.row.bin", d->test_row, 4 * nt) &&
              write_bin(b + "/R_test_coo.col.bin", d->test_col, 4 * nt);
    if (!ok) return fail(CDMF_ERR_IO, "cannot write binaries under " + b);
    FILE* f = std::fopen((b + "/meta_modified_all").c_str(), "w");
    if (!f) return fail(CDMF_ERR_IO, "cannot write meta_modified_all");
    std::fprintf(f, "%lld %lld\n%llu\n", (long long)d->m, (long long)d->n, (unsigned long long)nnz);
    std::fprintf(f, "R_train_coo.data.bin\nR_train_coo.row.bin\nR_train_coo.col.bin\n"
                    "R_train_csr.indptr.bin\nR_train_csr.indices.bin\nR_train_csr.data.bin\n"
                    "R_train_csc.indptr.bin\nR_train_csc.indices.bin\nR_train_csc.data.bin\n");
    std::fprintf(f, "%llu\nR_test_coo.data.bin\nR_test_coo.row.bin\nR_test_coo.col.bin\n", (unsigned long long)nt);
    std::fclose(f);
    f = std::fopen((b + "/meta").c_str(), "w");
    if (!f) return fail(CDMF_ERR_IO, "cannot write meta");
    std::fprintf(f, "%lld %lld\n%llu train.txt\n%llu test.txt\n", (long long)d->m, (long long)d->n,
                 (unsigned long long)nnz, (unsigned long long)nt);
    std::fclose(f);
    f = std::fopen((b + "/train.txt").c_str(), "w");
    if (!f) return fail(CDMF_ERR_IO, "cannot write train.txt");
    for (int64_t i = 0; i < d->m; ++i)
        for (uint64_t e = d->csr_ptr[i]; e < d->csr_ptr[i + 1]; ++e)
            std::fprintf(f, "%lld %u %.9g\n", (long long)i + 1, d->csr_idx[e] + 1, (double)d->csr_val[e]);
    std::fclose(f);
    f = std::fopen((b + "/test.txt").c_str(), "w");
    if (!f) return fail(CDMF_ERR_IO, "cannot write test.txt");
    for (uint64_t e = 0; e < nt; ++e)
        std::fprintf(f, "%u %u %.9g\n", d->test_row[e] + 1, d->test_col[e] + 1, (double)d->test_val[e]);
    std::fclose(f);
    return CDMF_OK;
}

}  // extern "C"
