/*
 * cdmf_data.h -- host-side data layer of the MI355X CCD++ solver.
 *
 * Plain C ABI (no C++ or torch types).  Everything here runs on the host:
 * the reader and writer of the reference's data-directory format and the
 * seeded synthetic generators used for the benchmark configurations.
 *
 * Reference interfaces these replace:
 *   cdmf_load_dir      load()            src/tools.cpp:271-347 and
 *                      SparseMatrix::read_binary_file / TestData::read_binary_file
 *                                        src/pmf_util.h:18-36, 80-108, 121-144
 *   cdmf_write_dir     the converters    data/simple/simple_compress.py:79-246
 *   cdmf_initial_col   initial_col()     src/tools.cpp:349-357
 *   cdmf_gen_*         the dataset generators data/{a,b,c}/{a,b,c}.py (here seeded and
 *                                        parameterised by the benchmark configurations)
 */
#ifndef CDMF_DATA_H
#define CDMF_DATA_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Training matrix in both compressed layouts plus the COO test set.
 * Mirrors SparseMatrix (src/pmf_util.h:13-114: u32 pointers and indices,
 * f32 values) and TestData (src/pmf_util.h:117-161).  Arrays are owned by
 * the library when filled by cdmf_gen_* / cdmf_load_dir; release them with
 * cdmf_dataset_free. */
typedef struct cdmf_dataset {
    int64_t m, n;
    uint64_t nnz, nnz_test;
    uint64_t max_row_nnz, max_col_nnz;
    uint32_t* csr_ptr; /* m+1 */
    uint32_t* csr_idx; /* nnz, column of each entry, ascending within a row */
    float* csr_val;    /* nnz */
    uint32_t* csc_ptr; /* n+1 */
    uint32_t* csc_idx; /* nnz, row of each entry, ascending within a column */
    float* csc_val;    /* nnz */
    uint32_t* test_row;
    uint32_t* test_col;
    float* test_val;
} cdmf_dataset;

/* Rating distributions of the generators. */
enum cdmf_rating_kind {
    CDMF_RATING_UNIFORM_1_5 = 0,  /* integers 1..5, equal probability */
    CDMF_RATING_NETFLIX_1_5 = 1,  /* integers 1..5 with P = .05 .10 .29 .34 .22 */
    CDMF_RATING_HALF_STAR = 2,    /* 0.5..5.0 in steps of 0.5, weighted toward 3.5-4.0 */
    CDMF_RATING_PLANTED_1_5 = 3   /* integers 1..5 with the NETFLIX_1_5 frequencies, ordered by a hidden
                                     rank-8 model: score = a_i + b_j + u_i.v_j + noise (Gaussian, seeded per
                                     row, column and entry), rating = the score's normal quantile mapped
                                     through the 1..5 frequencies (power-law generator only) */
};

/* Power-law synthetic ratings matrix.
 * Row degrees are drawn from a discrete power law P(d) ~ d^-row_alpha on
 * [1, max_row_degree] (row_alpha <= 0: the exponent is fitted so that the
 * mean degree equals (nnz_train+nnz_test)/m), then rescaled to the exact
 * total.  Each row picks distinct columns with probability proportional to
 * popularity rank^-col_exponent (ranks are a seeded permutation of the
 * columns).  nnz_test entries, chosen uniformly, are held out as the test
 * set.  Every random draw comes from xoshiro256** streams seeded from
 * `seed` and the row number, so the result does not depend on `threads`. */
typedef struct cdmf_gen_spec {
    int64_t m, n;
    uint64_t nnz_train, nnz_test;
    uint32_t max_row_degree; /* 0: n */
    double row_alpha;
    double col_exponent;
    int32_t rating_kind;
    int32_t threads;         /* 0: all hardware threads */
    uint64_t seed;
} cdmf_gen_spec;

/* Status codes shared with cdmf_hip.h. */
#define CDMF_OK 0
#define CDMF_ERR_INVALID (-1)
#define CDMF_ERR_IO (-6)
#define CDMF_ERR_ALLOC (-3)

int cdmf_gen_powerlaw(const cdmf_gen_spec* spec, cdmf_dataset* out);

/* nnz_train+nnz_test distinct positions drawn uniformly from the m x n grid
 * (Floyd's algorithm), ratings of `rating_kind`, a uniform hold-out of
 * nnz_test entries. */
int cdmf_gen_uniform(int64_t m, int64_t n, uint64_t nnz_train, uint64_t nnz_test, int32_t rating_kind,
                     uint64_t seed, cdmf_dataset* out);

/* The reference's CTest stress shapes (data/a/a.py:25-55, data/b/b.py:24-55,
 * data/c/c.py:25-43), seeded, with ratings uniform in 1..5:
 *   'a': 64,000 x 150; row 0 holds every column, every other row column 0
 *        and one column drawn from [1, n)            (128,148 train ratings)
 *   'b': 1,024,000 x 1,024,000; every row but the last holds one column
 *        drawn from [0, n-1) and column n-1, the last row every column
 *                                                    (3,071,998 train ratings)
 *   'c': 1,024,000 x 1,024,000; every row holds column 0 and one column
 *        drawn from [1, n)                           (2,048,000 train ratings)
 * The test set has train/100 entries, like the reference's, but at
 * positions drawn apart from the training ratings (the reference takes its
 * test lines from the head of the training file, so they overlap). */
int cdmf_gen_stress(char shape, uint64_t seed, cdmf_dataset* out);

/* Read a data directory (meta_modified_all + binaries).  A compressed
 * pointer file shorter than m+1 / n+1 entries is padded with nnz.  Errors:
 * CDMF_ERR_IO with a message available from cdmf_data_last_error(). */
int cdmf_load_dir(const char* dir, cdmf_dataset* out);

/* Write a data directory in the reference layout (binaries, meta_modified_all,
 * meta, train.txt, test.txt). */
int cdmf_write_dir(const char* dir, const cdmf_dataset* ds);

void cdmf_dataset_free(cdmf_dataset* ds);

/* Last error message of this thread's data-layer call. */
const char* cdmf_data_last_error(void);

/* initial_col of src/tools.cpp:349-357: X[t][i] = 0.1*rand()/RAND_MAX + 0.001
 * after srand(0), i-major draw order; X is rank-major [k][n]. */
void cdmf_initial_col(float* X, uint32_t k, uint32_t n);

/* X[e] = scale * U[0,1) for e < count, from a seeded xoshiro256** stream. */
void cdmf_init_uniform(float* X, uint64_t count, float scale, uint64_t seed);

#ifdef __cplusplus
}
#endif

#endif /* CDMF_DATA_H */
