// Host tools (see tools.h).  Each function states the reference behaviour it
// keeps; messages that scripts or users match on are kept verbatim.
#include "tools.h"

#include <cmath>
#include <cstring>
#include <string>

#include "cdmf_data.h"

namespace {
template <class T>
std::shared_ptr<T> own(T* p) {
    return std::shared_ptr<T>(p, [](T* q) { std::free(q); });
}
}  // namespace

void SparseMatrix::adopt(cdmf_dataset& d) {
    rows = (long)d.m;
    cols = (long)d.n;
    nnz = d.nnz;
    max_row_nnz = d.max_row_nnz;
    max_col_nnz = d.max_col_nnz;
    csr_row_ptr_ = own(d.csr_ptr);
    csr_col_indx_ = own(d.csr_idx);
    csr_val_ = own(d.csr_val);
    csc_col_ptr_ = own(d.csc_ptr);
    csc_row_indx_ = own(d.csc_idx);
    csc_val_ = own(d.csc_val);
    d.csr_ptr = d.csr_idx = d.csc_ptr = d.csc_idx = nullptr;
    d.csr_val = d.csc_val = nullptr;
}

// pmf_util.h:38-53: the transpose shares the arrays; CSR and CSC swap roles.
SparseMatrix SparseMatrix::get_shallow_transpose() const {
    SparseMatrix t;
    t.rows = cols;
    t.cols = rows;
    t.nnz = nnz;
    t.max_row_nnz = max_col_nnz;
    t.max_col_nnz = max_row_nnz;
    t.csr_row_ptr_ = csc_col_ptr_;
    t.csr_col_indx_ = csc_row_indx_;
    t.csr_val_ = csc_val_;
    t.csc_col_ptr_ = csr_row_ptr_;
    t.csc_row_indx_ = csr_col_indx_;
    t.csc_val_ = csr_val_;
    return t;
}

void TestData::adopt(cdmf_dataset& d) {
    rows = (unsigned long)d.m;
    cols = (unsigned long)d.n;
    nnz = d.nnz_test;
    test_row = own(d.test_row);
    test_col = own(d.test_col);
    test_val = own(d.test_val);
    d.test_row = d.test_col = nullptr;
    d.test_val = nullptr;
}

// tools.cpp:271-347: read meta_modified_all and the binaries; a missing meta
// file prints "Can't open meta input file." and exits with EXIT_FAILURE, a
// short or unreadable binary aborts.
void load(const char* srcdir, SparseMatrix& R, TestData& T) {
    cdmf_dataset d;
    auto t0 = std::chrono::high_resolution_clock::now();
    int rc = cdmf_load_dir(srcdir, &d);
    if (rc != CDMF_OK) {
        const char* why = cdmf_data_last_error();
        if (std::strcmp(why, "Can't open meta input file.") == 0) {
            printf("Can't open meta input file.\n");
            exit(EXIT_FAILURE);
        }
        fprintf(stderr, "%s\n", why);
        abort();
    }
    auto t1 = std::chrono::high_resolution_clock::now();
    R.adopt(d);
    T.adopt(d);
    std::chrono::duration<double> dt = t1 - t0;
    std::cout << "[info] Train TIMER: " << dt.count() << "s.\n";
    std::cout << "[info] Tests TIMER: " << 0.0 << "s.\n";
    cdmf_dataset_free(&d);
}

// tools.cpp:349-357
void initial_col(MatData& X, unsigned k, unsigned n) {
    std::vector<float> flat((size_t)k * n);
    cdmf_initial_col(flat.data(), k, n);
    X = MatData(k, VecData(n));
    for (unsigned t = 0; t < k; ++t) std::copy(flat.begin() + (size_t)t * n, flat.begin() + (size_t)(t + 1) * n, X[t].begin());
}

// tools.cpp:359-376 (help text; options of this build appended)
void exit_with_help() {
    printf("Usage: cdmf [options] data_dir\n"
           "options:\n"
           "    -c: path to the kernel code (accepted for compatibility; kernels are built ahead of time)\n"
           "    -k rank: set the rank (default 10)\n"
           "    -n threads: set the number of host threads (default 16)\n"
           "    -l lambda: set the regularization parameter lambda (default 0.1)\n"
           "    -t max_outer_iter: set the number of outer iterations (default 5)\n"
           "    -T max_inner_iter: set the number of inner iterations used in CCD (default 5)\n"
           "    -d device_type: select a device (0=gpu, 1=cpu, 2=mic) (default 0; only the HIP GPU exists)\n"
           "    -P platform_id: select the HIP device index (default 0)\n"
           "    -q verbose: show information or not (default 0)\n"
           "    -p predict: compute the test RMSE from the meta text files after the solve (default 0)\n"
           "    -r ref: compare W, H with W_ref.bin / H_ref.bin of the data directory (default 0)\n"
           "    -V version: 1 or 2 = degree-binned sweeps, 3 = tiled sweeps (default 1)\n"
           "    -nBlocks: number of blocks (accepted for compatibility)\n"
           "    -nThreadsPerBlock: number of threads per block (accepted for compatibility)\n"
           "    --order h|w: sweep H first (the reference) or W first (default h)\n"
           "    --precision fp32|bf16|fp64sums: storage of W and H (default fp32)\n"
           "    --graph 0|1: replay outer iterations from a captured HIP graph (default 0)\n"
           "    --zero-first 0|1: zero the first-swept factor at entry (default 1)\n");
    exit(EXIT_FAILURE);
}

// tools.cpp:378-465: "-x value" pairs, then the data directory.
parameter parse_command_line(int argc, char** argv) {
    parameter param{};
    int device_id = 0;
    int i;
    for (i = 1; i < argc; i++) {
        if (argv[i][0] != '-') break;
        if (++i >= argc) exit_with_help();
        const char* opt = argv[i - 1];
        const char* val = argv[i];
        if (std::strcmp(opt, "-nBlocks") == 0) {
            param.nBlocks = atoi(val);
        } else if (std::strcmp(opt, "-nThreadsPerBlock") == 0) {
            param.nThreadsPerBlock = atoi(val);
        } else if (std::strcmp(opt, "--order") == 0) {
            if (std::strcmp(val, "h") == 0) param.order = 0;
            else if (std::strcmp(val, "w") == 0) param.order = 1;
            else exit_with_help();
        } else if (std::strcmp(opt, "--precision") == 0) {
            if (std::strcmp(val, "fp32") == 0) param.precision = 0;
            else if (std::strcmp(val, "bf16") == 0) param.precision = 1;
            else if (std::strcmp(val, "fp64sums") == 0) param.precision = 2;
            else exit_with_help();
        } else if (std::strcmp(opt, "--graph") == 0) {
            param.use_graph = atoi(val);
        } else if (std::strcmp(opt, "--zero-first") == 0) {
            param.zero_first = atoi(val);
        } else {
            switch (opt[1]) {
                case 'c': snprintf(param.kcode_path, sizeof param.kcode_path, "%s", val); break;
                case 'k': param.k = (unsigned)atoi(val); break;
                case 'n': param.threads = atoi(val); break;
                case 'l': param.lambda = (VALUE_TYPE)atof(val); break;
                case 't': param.maxiter = atoi(val); break;
                case 'T': param.maxinneriter = atoi(val); break;
                case 'P': param.platform_id = (unsigned)atoi(val); break;
                case 'd': device_id = atoi(val); break;
                case 'p': param.do_predict = atoi(val); break;
                case 'r': param.do_ref = atoi(val); break;
                case 'q': param.verbose = atoi(val); break;
                case 'V': param.version = atoi(val); break;
                default:
                    fprintf(stderr, "unknown option: -%c\n", opt[1]);
                    exit_with_help();
            }
        }
    }
    switch (device_id) {
        case 0: snprintf(param.device_type, 4, "gpu"); break;
        case 1: snprintf(param.device_type, 4, "cpu"); break;
        case 2: snprintf(param.device_type, 4, "mic"); break;
        default:
            fprintf(stderr, "unknown device type!\n");
            exit_with_help();
    }
    printf("[info] - selected device type: %s, on platform with index: %u | Will be using %d threads | Value type is %s\n",
           param.device_type, param.platform_id, param.threads, sizeof(VALUE_TYPE) == 8 ? "double" : "float");
    if (i >= argc) exit_with_help();
    snprintf(param.scr_dir, sizeof param.scr_dir, "%s", argv[i]);
    return param;
}

// tools.cpp:467-487: entries off by more than 10 % of the reference value.
void golden_compare(MatData W, MatData W_ref, unsigned k, unsigned m) {
    unsigned bad = 0;
    for (unsigned t = 0; t < k; ++t)
        for (unsigned j = 0; j < m; ++j)
            if (std::fabs((double)W[t][j] - (double)W_ref[t][j]) > 0.1 * std::fabs((double)W_ref[t][j])) ++bad;
    if (bad == 0) {
        std::cout << "Check... PASS!" << std::endl;
    } else {
        unsigned entries = k * m;
        printf("Check... NO PASS! [%f%%] #Error = %u out of %u entries.\n", 100.0 * (double)bad / entries, bad, entries);
    }
}

// tools.cpp:489-541: RMSE over the 1-based "i j v" test file named in meta.
void calculate_rmse(const MatData& W_c, const MatData& H_c, const char* srcdir, const unsigned k) {
    std::string meta = std::string(srcdir) + "/meta";
    FILE* fp = fopen(meta.c_str(), "r");
    if (!fp) {
        printf("Can't open meta input file.\n");
        exit(EXIT_FAILURE);
    }
    char train_name[1024], test_name[1024];
    unsigned m, n, nnz, nnz_test;
    if (fscanf(fp, "%u %u", &m, &n) != 2 || fscanf(fp, "%u %1023s", &nnz, train_name) != 2 ||
        fscanf(fp, "%u %1023s", &nnz_test, test_name) != 2) {
        fprintf(stderr, "malformed meta file\n");
        abort();
    }
    fclose(fp);
    std::string test_file = std::string(srcdir) + "/" + test_name;
    FILE* tf = fopen(test_file.c_str(), "r");
    if (!tf) {
        printf("Can't open test file.\n");
        exit(EXIT_FAILURE);
    }
    double sum = 0.0;
    int count = 0, nans = 0;
    unsigned i, j;
    double v;
    while (fscanf(tf, "%u %u %lf", &i, &j, &v) == 3) {
        double pred = 0.0;
        if (i >= 1 && j >= 1 && i <= W_c[0].size() && j <= H_c[0].size())
            for (unsigned t = 0; t < k; ++t) pred += W_c[t][i - 1] * H_c[t][j - 1];
        else
            pred = NAN;
        double e2 = (pred - v) * (pred - v);
        if (std::isnan(e2)) ++nans;
        else sum += e2;
        ++count;
    }
    fclose(tf);
    if (count == 0) exit(EXIT_FAILURE);
    printf("[INFO] NaNs: [%lf%%], NaNs Count: %d out of %d entries.\n", (double)nans / count, nans, count);
    printf("[INFO] Test RMSE = %lf\n", std::sqrt(sum / count));
}

// tools.cpp:543-582 (host FP64 RMSE; the solver computes the same on the device).
double calculate_rmse_directly(MatData& W, MatData& H, TestData& T, unsigned rank, bool ifALS) {
    double sum = 0.0;
    unsigned long count = 0;
    for (unsigned long e = 0; e < T.nnz; ++e) {
        unsigned i = T.getTestRow()[e], j = T.getTestCol()[e];
        double pred = 0.0;
        for (unsigned t = 0; t < rank; ++t) pred += ifALS ? W[i][t] * H[j][t] : W[t][i] * H[t][j];
        double e2 = (pred - (double)T.getTestVal()[e]) * (pred - (double)T.getTestVal()[e]);
        if (!std::isnan(e2)) sum += e2;
        ++count;
    }
    if (count == 0) exit(EXIT_FAILURE);
    return std::sqrt(sum / count);
}

bool read_factor_file(const char* fname, MatData& X, unsigned k, long n) {
    FILE* f = fopen(fname, "rb");
    if (!f) return false;
    X = MatData(k, VecData(n));
    bool ok = true;
    for (unsigned t = 0; t < k && ok; ++t) ok = fread(X[t].data(), sizeof(VALUE_TYPE), (size_t)n, f) == (size_t)n;
    fclose(f);
    return ok;
}
