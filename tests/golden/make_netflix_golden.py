"""Write tests/golden/netflix_k40.npz: the oracle's answer for the full-size
solve, so that the GPU test of that solve needs no CPU solve of its own.

Workload: the seeded 480,189 x 17,770 matrix (99,072,112 train / 1,408,395
test ratings) of cdmf_amd.data.NETFLIX, k = 40, lambda = 0.05, 10 outer x 5
inner iterations, W from initial_col (src/tools.cpp:349-357), H = 0, H
swept first (src/cdmf_ref.cpp:47-140).

Two oracle runs (oracle/ccd_oracle.c, all host cores):
  * FP32 (or_ccd, the reference's arithmetic): test and train RMSE per outer
    iteration;
  * FP64 (or_ccd_f64): test and train RMSE per outer iteration, and W, H and
    the residual at the end, sampled: 4,096 seeded users (columns of W),
    4,096 seeded items (columns of H) and 65,536 seeded residual entries
    (CSR order), stored as float32.
Also stored: a SHA-256 of the generated matrix (every array of the data
set), so the GPU test can assert that it regenerated the same matrix, the
FP32 oracle's Frobenius distance to the FP64 one (whole W, H, residual),
and the wall time of each oracle solve.

usage: python tests/golden/make_netflix_golden.py   (about 10 minutes on 8 cores)
"""
from __future__ import annotations

import hashlib
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from cdmf_amd import data as D  # noqa: E402

SAMPLE_SEED = 20240
N_USERS, N_ITEMS, N_RES = 4096, 4096, 65536


def dataset_sha(ds) -> str:
    h = hashlib.sha256()
    for a in (ds.csr_ptr, ds.csr_idx, ds.csr_val, ds.csc_ptr, ds.csc_idx, ds.csc_val, ds.test_row, ds.test_col,
              ds.test_val):
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def samples(ds):
    rng = np.random.default_rng(SAMPLE_SEED)
    users = np.sort(rng.choice(ds.m, size=N_USERS, replace=False))
    items = np.sort(rng.choice(ds.n, size=N_ITEMS, replace=False))
    res = np.sort(rng.choice(ds.nnz, size=N_RES, replace=False))
    return users, items, res


def frob(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def main():
    cfg = D.NETFLIX_SOLVE
    t = time.perf_counter()
    ds = D.netflix()
    print("generated in %.1fs" % (time.perf_counter() - t), flush=True)
    k, lam, outer, inner = cfg.k, cfg.lam, cfg.outer, cfg.inner
    lam32 = float(np.float32(lam))
    W0 = D.initial_col(k, ds.m)
    H0 = np.zeros((k, ds.n), np.float32)
    threads = os.cpu_count() or 1
    out = dict(k=k, lam=lam, outer=outer, inner=inner, data_sha=dataset_sha(ds), threads=threads)

    tr32 = np.full(outer, np.nan)
    t = time.perf_counter()
    W32, H32, rm32, rv32, _ = oracle.ccd(ds, W0, H0, k, lam32, outer, inner, threads=threads, train_out=tr32)
    out["fp32_seconds"] = time.perf_counter() - t
    print("fp32 oracle %.1fs test %s train %s" % (out["fp32_seconds"], rm32, tr32), flush=True)
    out.update(fp32_test_rmse=rm32, fp32_train_rmse=tr32)

    tr64 = np.full(outer, np.nan)
    t = time.perf_counter()
    W64, H64, rm64, rv64, _ = oracle.ccd(ds, W0, H0, k, lam32, outer, inner, threads=threads, dtype="f64",
                                         train_out=tr64)
    out["fp64_seconds"] = time.perf_counter() - t
    print("fp64 oracle %.1fs test %s train %s" % (out["fp64_seconds"], rm64, tr64), flush=True)
    out.update(fp64_test_rmse=rm64, fp64_train_rmse=tr64)

    users, items, res = samples(ds)
    out.update(users=users.astype(np.int64), items=items.astype(np.int64), res_pos=res.astype(np.int64),
               fp64_W=W64[:, users].astype(np.float32), fp64_H=H64[:, items].astype(np.float32),
               fp64_res=rv64[res].astype(np.float32),
               fp64_W_rms=float(np.sqrt(np.mean(W64 ** 2))), fp64_H_rms=float(np.sqrt(np.mean(H64 ** 2))),
               fp64_res_rms=float(np.sqrt(np.mean(rv64 ** 2))),
               fp32_vs_fp64_frob=np.array([frob(W32, W64), frob(H32, H64), frob(rv32, rv64)]))
    print("fp32 vs fp64 Frobenius (W, H, residual):", out["fp32_vs_fp64_frob"], flush=True)
    np.savez_compressed(os.path.join(HERE, "netflix_k40.npz"), **out)
    print("wrote netflix_k40.npz", flush=True)


if __name__ == "__main__":
    main()
