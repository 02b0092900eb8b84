// Solver parameters: the fields and defaults of the reference's parameter
// class (src/pmf.h:4-26), plus the options of this implementation.
#ifndef CDMF_HOST_PMF_H
#define CDMF_HOST_PMF_H

#ifndef VALUE_TYPE
#define VALUE_TYPE float
#endif

class parameter {
public:
    int threads = 16;           // -n (used by the host loops only)
    int nBlocks = 1;            // -nBlocks (accepted; the work plan sizes launches itself)
    int nThreadsPerBlock = 32;  // -nThreadsPerBlock (accepted, see above)

    int do_ref = 0;      // -r: compare W, H with reference factors
    int do_predict = 0;  // -p: text-test RMSE after the solve

    int maxinneriter = 5;  // -T
    int maxiter = 5;       // -t
    unsigned k = 10;       // -k
    VALUE_TYPE lambda = (VALUE_TYPE)0.1;  // -l

    unsigned platform_id = 0;  // -P: HIP device index
    int verbose = 0;           // -q

    int version = 1;  // -V: 1 or 2 = binned sweeps, 3 = tiled sweeps

    char device_type[4] = {'g', 'p', 'u', '\0'};  // -d
    char scr_dir[1024] = "../data/simple";
    char kcode_path[1024] = "../kcode";  // -c: accepted, unused (kernels are compiled ahead of time)

    // extensions (long options)
    int order = 0;      // --order h|w
    int precision = 0;  // --precision fp32|bf16|fp64sums
    int use_graph = 0;  // --graph 0|1
    int zero_first = 1; // --zero-first 0|1
};

#endif
