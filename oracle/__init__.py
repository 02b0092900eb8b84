"""CPU oracle of the CCD++ hot path -- TEST INFRASTRUCTURE ONLY.

ctypes wrapper of oracle/_build/libccd_oracle.so (built from ccd_oracle.c by
oracle/Makefile).  Only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg may import this package, and only as the checker.

Function -> reference it restates (see ccd_oracle.c for the details):
  ccd        src/cdmf_ref.cpp:47-140     rmse        src/tools.cpp:543-582
  initial_col src/tools.cpp:349-357      dual_spmv   src/detail/kcode/csr5_spmv.cl:134-147
  ccd_ranks  slices of ccd (the CPU baseline); chained, equal to ccd bit for bit
  ccd_f64 / ccd_bf16 / objective / residual_f64 / sweep_f64 / train RMSE:
  no reference counterpart; pinned by the checks in tests/test_oracle.py
  (an independent numpy restatement of the BF16 twin, the twin built with
  identity rounding against ccd, mathematics for the rest).
  fold_rows / fold_features / mlp_forward / mlp_backward (foldin_oracle.c):
  the fold-in extension, no reference counterpart; pinned by
  tests/test_oracle_foldin.py (ridge solve, finite differences).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_build", "libccd_oracle.so")
LIB_BF16ID = os.path.join(HERE, "_build", "libccd_oracle_bf16id.so")
REF_HARNESS = os.path.join(HERE, "_ref", "ref_harness")

_lib = None
_lib_bf16id = None

u32p = C.POINTER(C.c_uint32)
f32p = C.POINTER(C.c_float)
f64p = C.POINTER(C.c_double)


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE, "oracle"], check=True)


def _declare(L: C.CDLL) -> None:
    i64, u64, u32, i32, flt, dbl = C.c_int64, C.c_uint64, C.c_uint32, C.c_int, C.c_float, C.c_double
    L.or_ccd.argtypes = [i64, i64, u32p, u32p, f32p, u32p, u32p, f32p, f32p, f32p, u32, flt, i32, i32, i32, i32,
                         i32, u64, u32p, u32p, f32p, f64p, f64p]
    L.or_ccd_bf16.argtypes = L.or_ccd.argtypes
    L.or_ccd_f64.argtypes = [i64, i64, u32p, u32p, f64p, u32p, u32p, f64p, f64p, f64p, u32, dbl, i32, i32, i32,
                             i32, i32, u64, u32p, u32p, f32p, f64p, f64p]


def lib_bf16_identity() -> C.CDLL:
    """The oracle built with the BF16 twin's rounding replaced by the identity."""
    global _lib_bf16id
    if _lib_bf16id is None:
        if not os.path.exists(LIB_BF16ID):
            build()
        L = C.CDLL(LIB_BF16ID)
        _declare(L)
        _lib_bf16id = L
    return _lib_bf16id


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        i64, u64, u32, i32, flt, dbl = C.c_int64, C.c_uint64, C.c_uint32, C.c_int, C.c_float, C.c_double
        _declare(L)
        L.or_ccd_ranks.argtypes = [i64, i64, u32p, u32p, f32p, u32p, u32p, f32p, f32p, f32p, u32, u32, flt, i32, i32,
                                   i32]
        L.or_rmse.argtypes = [f32p, f32p, i64, i64, u32, u64, u32p, u32p, f32p]
        L.or_rmse.restype = dbl
        L.or_rmse_f64.argtypes = [f64p, f64p, i64, i64, u32, u64, u32p, u32p, f32p]
        L.or_rmse_f64.restype = dbl
        L.or_initial_col.argtypes = [f32p, u32, u32]
        L.or_initial_col.restype = None
        L.or_bf16.argtypes = [flt]
        L.or_bf16.restype = flt
        L.or_dual_spmv.argtypes = [i64, u32p, u32p, f32p, f32p, f64p, f64p]
        L.or_dual_spmv.restype = None
        L.or_sweep_f64.argtypes = [i64, u32p, u32p, f32p, f32p, flt, f64p]
        L.or_sweep_f64.restype = None
        L.or_objective.argtypes = [i64, i64, u32p, f32p, u32p, f32p, f32p, u32, dbl]
        L.or_objective.restype = dbl
        L.or_residual_f64.argtypes = [i64, i64, u32p, u32p, f32p, f32p, f32p, u32, f64p]
        L.or_residual_f64.restype = None
        L.or_fold_rows.argtypes = [i64, u32p, u32p, f32p, u32, i64, f32p, flt, i32, f32p, i32]
        L.or_fold_rows.restype = i32
        L.or_fold_features.argtypes = [i64, u32p, u32p, f32p, u32, i64, f32p, f32p]
        L.or_fold_features.restype = None
        L.or_mlp_forward.argtypes = [i64, i32, i32, i32, f64p, f64p, f64p]
        L.or_mlp_forward.restype = None
        L.or_mlp_backward.argtypes = [i64, i32, i32, i32, f64p, f64p, f64p, f64p]
        L.or_mlp_backward.restype = dbl
        _lib = L
    return _lib


def _p(a, t):
    return a.ctypes.data_as(C.POINTER(t))


def _c(a, dtype):
    return np.ascontiguousarray(a, dtype=dtype)


def ccd(ds, W, H, k, lam, outer, inner, order=0, zero_first=True, threads=0, dtype="f32", train_out=None,
        _lib=None):
    """Run the restated cdmf_ref on copies of ds's values and of W, H.

    Returns (W, H, rmse_per_iter, csr_residual, csc_residual).  dtype is
    "f32" (the reference), "bf16" (BF16-stored factors) or "f64".
    train_out: a float64 array of at least `outer` entries that receives the
    train RMSE after every outer iteration."""
    L = _lib or lib()
    if train_out is not None:
        assert train_out.dtype == np.float64 and train_out.flags.c_contiguous and train_out.size >= outer
    tp = train_out.ctypes.data_as(f64p) if train_out is not None else None
    cp, ci, cq, ki = _c(ds.csr_ptr, np.uint32), _c(ds.csr_idx, np.uint32), _c(ds.csc_ptr, np.uint32), _c(ds.csc_idx, np.uint32)
    tr, tc, tv = _c(ds.test_row, np.uint32), _c(ds.test_col, np.uint32), _c(ds.test_val, np.float32)
    rm = np.full(max(outer, 1), np.nan)
    if dtype == "f64":
        rv, kv = ds.csr_val.astype(np.float64), ds.csc_val.astype(np.float64)
        Wc, Hc = np.array(W, dtype=np.float64), np.array(H, dtype=np.float64)
        rc = L.or_ccd_f64(ds.m, ds.n, _p(cp, C.c_uint32), _p(ci, C.c_uint32), _p(rv, C.c_double),
                          _p(cq, C.c_uint32), _p(ki, C.c_uint32), _p(kv, C.c_double), _p(Wc, C.c_double),
                          _p(Hc, C.c_double), k, lam, outer, inner, order, int(zero_first), threads, tr.size,
                          _p(tr, C.c_uint32), _p(tc, C.c_uint32), _p(tv, C.c_float), _p(rm, C.c_double), tp)
    else:
        rv, kv = ds.csr_val.astype(np.float32).copy(), ds.csc_val.astype(np.float32).copy()
        Wc, Hc = np.array(W, dtype=np.float32), np.array(H, dtype=np.float32)
        f = L.or_ccd if dtype == "f32" else L.or_ccd_bf16
        rc = f(ds.m, ds.n, _p(cp, C.c_uint32), _p(ci, C.c_uint32), _p(rv, C.c_float),
               _p(cq, C.c_uint32), _p(ki, C.c_uint32), _p(kv, C.c_float), _p(Wc, C.c_float), _p(Hc, C.c_float),
               k, lam, outer, inner, order, int(zero_first), threads, tr.size,
               _p(tr, C.c_uint32), _p(tc, C.c_uint32), _p(tv, C.c_float), _p(rm, C.c_double), tp)
    if rc != 0:
        raise RuntimeError("oracle failed (%d)" % rc)
    return Wc, Hc, rm[:outer], rv, kv


class RankState:
    """Arrays of an H-first solve advanced slice by slice with ccd_ranks:
    contiguous copies of the index arrays, the two residual copies (R at
    the start) and W, H, all updated in place.  Building it copies; the
    rank slices do not."""

    def __init__(self, ds, W, H):
        self.m, self.n = ds.m, ds.n
        self.cp, self.ci, self.cq, self.ki = (_c(a, np.uint32) for a in (ds.csr_ptr, ds.csr_idx, ds.csc_ptr,
                                                                        ds.csc_idx))
        self.rv, self.kv = ds.csr_val.astype(np.float32).copy(), ds.csc_val.astype(np.float32).copy()
        self.W, self.H = np.array(W, dtype=np.float32), np.array(H, dtype=np.float32)


def ccd_ranks(st: RankState, t0, t1, lam, oiter, inner, threads=0) -> None:
    """Ranks [t0, t1) of outer iteration `oiter` of an H-first solve, in
    place on `st` (the caller zeroes H before outer iteration 1, as or_ccd
    does at entry)."""
    L = lib()
    rc = L.or_ccd_ranks(st.m, st.n, _p(st.cp, C.c_uint32), _p(st.ci, C.c_uint32), _p(st.rv, C.c_float),
                        _p(st.cq, C.c_uint32), _p(st.ki, C.c_uint32), _p(st.kv, C.c_float), _p(st.W, C.c_float),
                        _p(st.H, C.c_float), t0, t1, lam, oiter, inner, threads)
    if rc != 0:
        raise RuntimeError("oracle failed (%d)" % rc)


def rmse(W, H, ds) -> float:
    L = lib()
    W, H = _c(W, np.float32), _c(H, np.float32)
    tr, tc, tv = _c(ds.test_row, np.uint32), _c(ds.test_col, np.uint32), _c(ds.test_val, np.float32)
    return L.or_rmse(_p(W, C.c_float), _p(H, C.c_float), ds.m, ds.n, W.shape[0], tr.size, _p(tr, C.c_uint32),
                     _p(tc, C.c_uint32), _p(tv, C.c_float))


def initial_col(k: int, n: int) -> np.ndarray:
    X = np.empty((k, n), dtype=np.float32)
    lib().or_initial_col(_p(X, C.c_float), k, n)
    return X


def bf16(x: float) -> float:
    return lib().or_bf16(x)


def dual_spmv(rows, ptr, idx, val, x):
    L = lib()
    ptr, idx, val, x = _c(ptr, np.uint32), _c(idx, np.uint32), _c(val, np.float32), _c(x, np.float32)
    yt, yb = np.empty(rows), np.empty(rows)
    L.or_dual_spmv(rows, _p(ptr, C.c_uint32), _p(idx, C.c_uint32), _p(val, C.c_float), _p(x, C.c_float),
                   _p(yt, C.c_double), _p(yb, C.c_double))
    return yt, yb


def sweep_f64(rows, ptr, idx, val, x, lam):
    L = lib()
    ptr, idx, val, x = _c(ptr, np.uint32), _c(idx, np.uint32), _c(val, np.float32), _c(x, np.float32)
    y = np.empty(rows)
    L.or_sweep_f64(rows, _p(ptr, C.c_uint32), _p(idx, C.c_uint32), _p(val, C.c_float), _p(x, C.c_float), lam,
                   _p(y, C.c_double))
    return y


def objective(ds, csr_res, W, H, lam) -> float:
    L = lib()
    cp, cq = _c(ds.csr_ptr, np.uint32), _c(ds.csc_ptr, np.uint32)
    r, W, H = _c(csr_res, np.float32), _c(W, np.float32), _c(H, np.float32)
    return L.or_objective(ds.m, ds.n, _p(cp, C.c_uint32), _p(r, C.c_float), _p(cq, C.c_uint32), _p(W, C.c_float),
                          _p(H, C.c_float), W.shape[0], lam)


def residual_f64(ds, W, H):
    L = lib()
    cp, ci, v = _c(ds.csr_ptr, np.uint32), _c(ds.csr_idx, np.uint32), _c(ds.csr_val, np.float32)
    W, H = _c(W, np.float32), _c(H, np.float32)
    out = np.empty(ds.nnz)
    L.or_residual_f64(ds.m, ds.n, _p(cp, C.c_uint32), _p(ci, C.c_uint32), _p(v, C.c_float), _p(W, C.c_float),
                      _p(H, C.c_float), W.shape[0], _p(out, C.c_double))
    return out


def fold_rows(ptr, idx, val, H, lam, sweeps, W0=None, threads=0):
    """CCD++ fold-in of CSR rows against fixed item factors H [k][n]:
    returns W [rows][k] (float32) after `sweeps` sweeps from W0 (zeros)."""
    L = lib()
    ptr, idx, val, H = _c(ptr, np.uint32), _c(idx, np.uint32), _c(val, np.float32), _c(H, np.float32)
    k, n = H.shape
    rows = ptr.size - 1
    W = np.zeros((rows, k), np.float32) if W0 is None else np.array(W0, dtype=np.float32, copy=True)
    rc = L.or_fold_rows(rows, _p(ptr, C.c_uint32), _p(idx, C.c_uint32), _p(val, C.c_float), k, n, _p(H, C.c_float),
                        lam, sweeps, _p(W, C.c_float), threads)
    if rc:
        raise MemoryError("or_fold_rows")
    return W


def fold_features(ptr, idx, val, H):
    """Fold-in features [rows][2k + 2]: weighted mean and variance of the rated
    items' H columns, log(1 + degree), mean rating."""
    L = lib()
    ptr, idx, val, H = _c(ptr, np.uint32), _c(idx, np.uint32), _c(val, np.float32), _c(H, np.float32)
    k, n = H.shape
    rows = ptr.size - 1
    F = np.empty((rows, 2 * k + 2), np.float32)
    L.or_fold_features(rows, _p(ptr, C.c_uint32), _p(idx, C.c_uint32), _p(val, C.c_float), k, n, _p(H, C.c_float),
                       _p(F, C.c_float))
    return F


def mlp_forward(params, X, dims):
    """FP64 MLP forward; params flat (W1, b1, W2, b2, W3, b3), dims (in, hid, out)."""
    L = lib()
    P, X = _c(params, np.float64), _c(X, np.float64)
    nin, hid, out = dims
    Y = np.empty((X.shape[0], out))
    L.or_mlp_forward(X.shape[0], nin, hid, out, _p(P, C.c_double), _p(X, C.c_double), _p(Y, C.c_double))
    return Y


def mlp_backward(params, X, T, dims):
    """FP64 (loss, gradient) of mean((MLP(X) - T)^2) w.r.t. the flat parameters."""
    L = lib()
    P, X, T = _c(params, np.float64), _c(X, np.float64), _c(T, np.float64)
    nin, hid, out = dims
    G = np.empty_like(P)
    loss = L.or_mlp_backward(X.shape[0], nin, hid, out, _p(P, C.c_double), _p(X, C.c_double), _p(T, C.c_double),
                             _p(G, C.c_double))
    return loss, G
