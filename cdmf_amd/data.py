"""Synthetic data sets of the benchmark configurations and host-side init.

Every generator is seeded; the same seed gives the same matrix on any
machine and any thread count (xoshiro256** streams per row, see
cdmf_amd/csrc/data/data.cpp).
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass

import numpy as np

from . import _native as N
from .datadir import Dataset


@dataclass(frozen=True)
class SolveConfig:
    """Hyper-parameters of one CCD++ run (the CLI's -k -l -t -T)."""

    k: int
    lam: float
    outer: int
    inner: int


# The small CPU-runnable case: 2,000 x 1,000, 50,000 distinct uniform
# positions, integer ratings 1..5 uniform, 10 % held out; k=10, lambda=0.05,
# 5 outer x 5 inner.  W = 0 and H ~ U[0, 1/sqrt(k)), so it runs W first.
SMALL = dict(m=2000, n=1000, nnz_train=45000, nnz_test=5000, seed=0)
SMALL_SOLVE = SolveConfig(k=10, lam=0.05, outer=5, inner=5)

# The large case: 480,189 x 17,770 with 99,072,112 train + 1,408,395 test
# ratings, power-law row degrees (min 1, max 17,000), column popularity
# rank^-1.0, ratings 1..5 with P = [.05 .10 .29 .34 .22]; k=40, lambda=0.05,
# 10 outer x 5 inner.  The row-degree exponent is fitted (row_alpha = 0) so
# that the law on [1, 17,000] has the matrix's mean degree: about 1.38
# (row_degree_exponent).  A Zipf exponent of 1.2 on the same range would
# give a mean of about 620 ratings per user, three times the target;
# rescaling its draws to the total would cut the largest degrees to about
# 5,800.
NETFLIX = dict(m=480189, n=17770, nnz_train=99072112, nnz_test=1408395, max_row_degree=17000,
               row_alpha=0.0, col_exponent=1.0, rating_kind=1, seed=1)
NETFLIX_SOLVE = SolveConfig(k=40, lam=0.05, outer=10, inner=5)

# The held-out-users fold-in study on the large shape: the same matrix, and
# the same shape with ratings planted from a hidden rank-8 model (same 1..5
# frequencies), on which a warm start has something to learn.
NETFLIX_PLANTED = dict(NETFLIX, rating_kind=3)
FOLDIN_SOLVE = NETFLIX_SOLVE

# The precision-study case: 138,493 x 26,744 with 20.0M ratings (10 % test),
# power-law degrees with exponent 1.1 on the column side, half-star ratings.
ML20M = dict(m=138493, n=26744, nnz_train=18000237, nnz_test=2000026, max_row_degree=9254,
             row_alpha=0.0, col_exponent=1.1, rating_kind=2, seed=2)
ML20M_SOLVE = SolveConfig(k=64, lam=0.05, outer=10, inner=5)


# The batched case: 64 independent problems, rows log-uniform in [1k, 200k],
# columns log-uniform in [500, 20k], ratings log-uniform in [10k, 5M]
# (power-law degrees, integer ratings 1..5, 10 % held out), k from
# {8, 16, 32, 64}, lambda from {0.01, 0.05, 0.1}, 8 outer x 5 inner.  Two
# adjustments keep every draw a valid power-law matrix: at most half of the
# cells are rated, and a problem with fewer than two ratings per row keeps
# only nnz / 2 rows.
BATCH = dict(count=64, rows=(1000, 200000), cols=(500, 20000), nnz=(10000, 5000000), ks=(8, 16, 32, 64),
             lambdas=(0.01, 0.05, 0.1), outer=8, inner=5, seed=4)


@dataclass(frozen=True)
class BatchProblem:
    m: int
    n: int
    nnz: int  # train + test
    seed: int
    solve: SolveConfig


def batch_specs(count: int = BATCH["count"], seed: int = BATCH["seed"], scale: float = 1.0) -> list:
    """The shapes and hyper-parameters of the batched case (`scale` shrinks
    rows, columns and ratings for quick checks)."""
    rng = np.random.default_rng(seed)

    def logu(lo, hi):
        return math.exp(rng.uniform(math.log(lo), math.log(hi)))

    out = []
    for i in range(count):
        m = max(8, int(logu(*BATCH["rows"]) * scale))
        n = max(8, int(logu(*BATCH["cols"]) * scale))
        nnz = max(64, int(logu(*BATCH["nnz"]) * scale))
        nnz = min(nnz, (m * n) // 2)
        if nnz < 2 * m:
            m = max(8, nnz // 2)
            nnz = min(nnz, (m * n) // 2)
        k = int(rng.choice(BATCH["ks"]))
        lam = float(rng.choice(BATCH["lambdas"]))
        out.append(BatchProblem(m, n, nnz, seed * 1000 + i, SolveConfig(k, lam, BATCH["outer"], BATCH["inner"])))
    return out


def batch_dataset(bp: BatchProblem, threads: int = 0) -> Dataset:
    test = bp.nnz // 10
    return gen_powerlaw(bp.m, bp.n, bp.nnz - test, test, max_row_degree=bp.n, col_exponent=1.0, rating_kind=0,
                        seed=bp.seed, threads=threads)


def gen_uniform(m: int, n: int, nnz_train: int, nnz_test: int, seed: int = 0,
                rating_kind: int = 0) -> Dataset:
    cd = N.CDataset()
    N.check_data(N.data_lib().cdmf_gen_uniform(m, n, nnz_train, nnz_test, rating_kind, seed, C.byref(cd)))
    return N.dataset_from_c(cd)


# The reference's CTest stress shapes and their flags (CMakeLists.txt:61-63:
# -k 10 -t 5 -T 3 -l 0.05).
STRESS_SHAPES = ("a", "b", "c")
STRESS_SOLVE = SolveConfig(k=10, lam=0.05, outer=5, inner=3)


def gen_stress(shape: str, seed: int = 0) -> Dataset:
    """data/{a,b,c}/{a,b,c}.py of the reference, seeded (cdmf_gen_stress):
    a dense first row (a), a dense last row and last column (b) or a dense
    first column (c), every other row holding two ratings; the test set
    (train / 100 entries) lies apart from the training ratings."""
    cd = N.CDataset()
    N.check_data(N.data_lib().cdmf_gen_stress(shape.encode(), seed, C.byref(cd)))
    return N.dataset_from_c(cd)


def gen_powerlaw(m: int, n: int, nnz_train: int, nnz_test: int, max_row_degree: int = 0,
                 row_alpha: float = 0.0, col_exponent: float = 1.0, rating_kind: int = 1,
                 seed: int = 0, threads: int = 0) -> Dataset:
    spec = N.CGenSpec(m=m, n=n, nnz_train=nnz_train, nnz_test=nnz_test, max_row_degree=max_row_degree,
                      row_alpha=row_alpha, col_exponent=col_exponent, rating_kind=rating_kind,
                      threads=threads, seed=seed)
    cd = N.CDataset()
    N.check_data(N.data_lib().cdmf_gen_powerlaw(C.byref(spec), C.byref(cd)))
    return N.dataset_from_c(cd)


def row_degree_exponent(m: int, nnz_train: int, nnz_test: int, max_row_degree: int, n: int,
                        row_alpha: float = 0.0, **_) -> float:
    """The row-degree exponent the power-law generator uses for a spec: the
    given row_alpha, or, for row_alpha = 0, the exponent a whose law
    P(d) ~ d^-a on [1, max_row_degree] has mean degree (train + test) / m
    (the bisection of cdmf_gen_powerlaw, restated)."""
    if row_alpha > 0.0:
        return row_alpha
    D = min(max_row_degree or n, n)
    d = np.arange(1, D + 1, dtype=np.float64)
    target = (nnz_train + nnz_test) / m

    def mean(a):
        w = d ** -a
        return float((w * d).sum() / w.sum())

    if mean(0.0) < target:
        return 0.0
    lo, hi = 0.0, 6.0
    for _ in range(60):
        mid = 0.5 * (lo + hi)
        if mean(mid) > target:
            lo = mid
        else:
            hi = mid
    return 0.5 * (lo + hi)


def small() -> Dataset:
    return gen_uniform(**SMALL)


def netflix(threads: int = 0) -> Dataset:
    return gen_powerlaw(threads=threads, **NETFLIX)


def netflix_planted(threads: int = 0) -> Dataset:
    return gen_powerlaw(threads=threads, **NETFLIX_PLANTED)


def ml20m(threads: int = 0) -> Dataset:
    return gen_powerlaw(threads=threads, **ML20M)


def load_dir(path: str) -> Dataset:
    cd = N.CDataset()
    N.check_data(N.data_lib().cdmf_load_dir(path.encode(), C.byref(cd)))
    return N.dataset_from_c(cd)


def write_dir(path: str, ds: Dataset) -> None:
    import os
    os.makedirs(path, exist_ok=True)
    cd = N.dataset_to_c(ds)
    N.check_data(N.data_lib().cdmf_write_dir(path.encode(), C.byref(cd)))


def initial_col(k: int, n: int) -> np.ndarray:
    """src/tools.cpp:349-357: rank-major [k][n] float32 init (srand(0), glibc rand)."""
    X = np.empty((k, n), dtype=np.float32)
    N.data_lib().cdmf_initial_col(N.ptr(X, C.c_float), k, n)
    return X


def init_uniform(shape, scale: float, seed: int) -> np.ndarray:
    X = np.empty(shape, dtype=np.float32)
    N.data_lib().cdmf_init_uniform(N.ptr(X, C.c_float), X.size, scale, seed)
    return X


def small_init(ds: Dataset, k: int = SMALL_SOLVE.k, seed: int = 0):
    """W = 0 and H ~ U[0, 1/sqrt(k)) for the small case."""
    W = np.zeros((k, ds.m), dtype=np.float32)
    H = init_uniform((k, ds.n), 1.0 / math.sqrt(k), seed)
    return W, H
