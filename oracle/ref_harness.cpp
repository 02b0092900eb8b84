// Golden-vector harness around the reference's own OpenMP solver.
//
// TEST INFRASTRUCTURE ONLY.  Built by oracle/Makefile together with the
// reference's src/cdmf_ref.cpp and src/tools.cpp, compiled where they lie
// under the reference checkout, into oracle/_ref/.  It is this project's
// code calling the reference's functions unchanged:
//   load()                   src/tools.cpp:271-347
//   initial_col()            src/tools.cpp:349-357
//   cdmf_ref()               src/cdmf_ref.cpp:47-140
//   calculate_rmse_directly  src/tools.cpp:543-582
//
// usage: ref_harness <data_dir> <out_dir> <k> <lambda> <maxiter> <maxinner> [W_init.bin [H_init.bin]]
//
// For it = 1..maxiter it reloads the data, runs cdmf_ref with maxiter = it
// from the same initial factors, and writes
//   <out_dir>/W_<it>.bin, H_<it>.bin   float32 rank-major [k][m] / [k][n]
//   <out_dir>/rmse.txt                 one %.17g line per outer iteration
// and, for the last it, the residual left in R:
//   <out_dir>/res_csr.bin, res_csc.bin float32 [nnz]
// The initial factors default to initial_col() (as src/main.cpp:28-31 does);
// W_init.bin / H_init.bin (float32 [k][m] / [k][n]) override them.
#include "tools.h"
#include "cdmf.h"

std::chrono::duration<double> deltaT12;
std::chrono::duration<double> deltaTAB;

static bool read_mat(const char* fname, MatData& X, unsigned k, long n) {
    FILE* f = fopen(fname, "rb");
    if (!f) return false;
    X = MatData(k, VecData(n));
    for (unsigned t = 0; t < k; ++t) {
        if (fread(X[t].data(), sizeof(float), (size_t)n, f) != (size_t)n) { fclose(f); return false; }
    }
    fclose(f);
    return true;
}

static void write_mat(const std::string& fname, const MatData& X, unsigned k, long n) {
    FILE* f = fopen(fname.c_str(), "wb");
    if (!f) { fprintf(stderr, "cannot write %s\n", fname.c_str()); exit(2); }
    for (unsigned t = 0; t < k; ++t) fwrite(X[t].data(), sizeof(float), (size_t)n, f);
    fclose(f);
}

int main(int argc, char** argv) {
    if (argc < 7) {
        fprintf(stderr, "usage: %s data_dir out_dir k lambda maxiter maxinner [W_init.bin [H_init.bin]]\n", argv[0]);
        return 2;
    }
    const char* dir = argv[1];
    std::string out = argv[2];
    parameter param{};
    param.k = (unsigned)atoi(argv[3]);
    param.lambda = (VALUE_TYPE)atof(argv[4]);
    int maxiter = atoi(argv[5]);
    param.maxinneriter = atoi(argv[6]);
    param.threads = 4;

    FILE* rf = fopen((out + "/rmse.txt").c_str(), "w");
    if (!rf) { fprintf(stderr, "cannot write rmse.txt\n"); return 2; }
    for (int it = 1; it <= maxiter; ++it) {
        SparseMatrix R;
        TestData T;
        load(dir, R, T);
        MatData W, H;
        initial_col(W, param.k, R.rows);
        initial_col(H, param.k, R.cols);
        if (argc > 7 && !read_mat(argv[7], W, param.k, R.rows)) { fprintf(stderr, "bad W_init\n"); return 2; }
        if (argc > 8 && !read_mat(argv[8], H, param.k, R.cols)) { fprintf(stderr, "bad H_init\n"); return 2; }
        param.maxiter = it;
        cdmf_ref(R, W, H, T, param);
        double rmse = calculate_rmse_directly(W, H, T, param.k, false);
        fprintf(rf, "%.17g\n", rmse);
        write_mat(out + "/W_" + std::to_string(it) + ".bin", W, param.k, R.rows);
        write_mat(out + "/H_" + std::to_string(it) + ".bin", H, param.k, R.cols);
        if (it == maxiter) {
            FILE* f = fopen((out + "/res_csr.bin").c_str(), "wb");
            fwrite(R.get_csr_val(), sizeof(float), R.nnz, f);
            fclose(f);
            f = fopen((out + "/res_csc.bin").c_str(), "wb");
            fwrite(R.get_csc_val(), sizeof(float), R.nnz, f);
            fclose(f);
        }
    }
    fclose(rf);
    return 0;
}
