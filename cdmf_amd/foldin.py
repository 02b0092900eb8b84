"""Fold-in of new users with a learned warm start (include/cdmf_foldin.h).

A new user's factor w_i is found with the item factors H fixed by CCD++
sweeps over the user's ratings (plain fold-in: from w_i = 0).  The warm
start replaces the zero by the output of a small MLP trained on the
solver's own (features of a user's ratings, converged w_i) pairs:
  features  rating-weighted mean and variance of the rated items' H
            columns (2k), log(1 + degree), mean rating
  network   2k + 2 -> 256 -> 256 -> k, GELU, trained in FP32 with Adam,
            run with bfloat16 weights
All of it runs on the GPU (kernels/foldin_kernels.hip, mlp_kernels.hip).
`experiment` reproduces the held-out-users study: solve on 90 % of the
users, train the network on them, fold in the other 10 % both ways.
"""
from __future__ import annotations

import ctypes as C
import math
import time
from dataclasses import dataclass, field

import numpy as np

from . import _native as N
from .datadir import Dataset
from .solver import CCDError, Solver, VARIANT_TILED, hip_lib

_sig_done = False


def _lib():
    global _sig_done
    lib = hip_lib()
    if not _sig_done:
        vp, u32p, f32p, f64p = C.c_void_p, N.u32p, N.f32p, N.f64p
        sig = {
            "ccd_mlp_create": ([C.c_int32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint64, C.POINTER(vp)], C.c_int32),
            "ccd_mlp_destroy": ([vp], C.c_int32),
            "ccd_mlp_last_error": ([vp], C.c_char_p),
            "ccd_mlp_param_count": ([vp], C.c_uint64),
            "ccd_mlp_get_params": ([vp, f32p], C.c_int32),
            "ccd_mlp_set_params": ([vp, f32p], C.c_int32),
            "ccd_mlp_set_norm": ([vp, f32p, f32p, f32p, f32p], C.c_int32),
            "ccd_mlp_forward": ([vp, C.c_uint64, f32p, f32p, C.c_int32], C.c_int32),
            "ccd_mlp_loss_grad": ([vp, C.c_uint64, f32p, f32p, f64p, f32p], C.c_int32),
            "ccd_mlp_train": ([vp, C.c_uint64, f32p, f32p, C.c_uint32, C.c_uint32, C.c_float, C.c_uint64, f64p],
                              C.c_int32),
            "ccd_fold_create": ([C.c_int32, C.c_uint32, C.c_int64, f32p, C.c_float, C.POINTER(vp)], C.c_int32),
            "ccd_fold_destroy": ([vp], C.c_int32),
            "ccd_fold_last_error": ([vp], C.c_char_p),
            "ccd_fold_features": ([vp, C.c_int64, u32p, u32p, f32p, f32p], C.c_int32),
            "ccd_fold_rows": ([vp, C.c_int64, u32p, u32p, f32p, f32p, C.c_int32, f32p, f64p], C.c_int32),
            "ccd_fold_rows_noalloc": ([vp, C.c_int64, u32p, u32p, f32p, f32p, C.c_int32, f32p, f64p], C.c_int32),
            "ccd_fold_warm": ([vp, vp, C.c_int64, u32p, u32p, f32p, C.c_int32, C.c_int32, f32p, f64p], C.c_int32),
        }
        for name, (args, res) in sig.items():
            f = getattr(lib, name)
            f.argtypes = args
            f.restype = res
        _sig_done = True
    return lib


FOLDIN_SYMBOLS = ("ccd_mlp_create", "ccd_mlp_destroy", "ccd_mlp_last_error", "ccd_mlp_param_count",
                  "ccd_mlp_get_params", "ccd_mlp_set_params", "ccd_mlp_set_norm", "ccd_mlp_forward",
                  "ccd_mlp_loss_grad", "ccd_mlp_train", "ccd_fold_create", "ccd_fold_destroy",
                  "ccd_fold_last_error", "ccd_fold_features", "ccd_fold_rows", "ccd_fold_rows_noalloc",
                  "ccd_fold_warm")


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _u32(a):
    return np.ascontiguousarray(a, dtype=np.uint32)


class MLP:
    """in -> hidden -> hidden -> out with GELU, parameters flat
    (W1 [hid][in], b1, W2 [hid][hid], b2, W3 [out][hid], b3)."""

    def __init__(self, nin: int, hidden: int, nout: int, seed: int = 0, device: int = 0):
        self._lib = _lib()
        self._m = C.c_void_p()
        rc = self._lib.ccd_mlp_create(device, nin, hidden, nout, seed, C.byref(self._m))
        if rc != 0:
            raise CCDError(rc, "ccd_mlp_create failed")
        self.dims = (nin, hidden, nout)

    def _check(self, rc):
        if rc < 0:
            raise CCDError(rc, self._lib.ccd_mlp_last_error(self._m).decode())

    def close(self):
        if self._m:
            self._lib.ccd_mlp_destroy(self._m)
            self._m = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def param_count(self) -> int:
        return int(self._lib.ccd_mlp_param_count(self._m))

    def get_params(self) -> np.ndarray:
        P = np.empty(self.param_count, np.float32)
        self._check(self._lib.ccd_mlp_get_params(self._m, N.ptr(P, C.c_float)))
        return P

    def set_params(self, P) -> None:
        P = _f32(P)
        assert P.size == self.param_count
        self._check(self._lib.ccd_mlp_set_params(self._m, N.ptr(P, C.c_float)))

    def set_norm(self, in_mean, in_std, out_mean, out_std) -> None:
        a, b, c, d = (_f32(x) for x in (in_mean, in_std, out_mean, out_std))
        self._check(self._lib.ccd_mlp_set_norm(self._m, N.ptr(a, C.c_float), N.ptr(b, C.c_float),
                                               N.ptr(c, C.c_float), N.ptr(d, C.c_float)))

    def forward(self, X, bf16: int = 0) -> np.ndarray:
        """bf16: 0 FP32, 1 bfloat16 weights, 2 matrix cores (ccd_mlp_forward)."""
        X = _f32(X)
        Y = np.empty((X.shape[0], self.dims[2]), np.float32)
        self._check(self._lib.ccd_mlp_forward(self._m, X.shape[0], N.ptr(X, C.c_float), N.ptr(Y, C.c_float), int(bf16)))
        return Y

    def loss_grad(self, Xn, Tn):
        Xn, Tn = _f32(Xn), _f32(Tn)
        loss = C.c_double()
        G = np.empty(self.param_count, np.float32)
        self._check(self._lib.ccd_mlp_loss_grad(self._m, Xn.shape[0], N.ptr(Xn, C.c_float), N.ptr(Tn, C.c_float),
                                                C.byref(loss), N.ptr(G, C.c_float)))
        return loss.value, G

    def train(self, X, T, epochs: int, batch: int = 1024, lr: float = 1e-3, seed: int = 0) -> float:
        X, T = _f32(X), _f32(T)
        loss = C.c_double()
        self._check(self._lib.ccd_mlp_train(self._m, X.shape[0], N.ptr(X, C.c_float), N.ptr(T, C.c_float), epochs,
                                            batch, lr, seed, C.byref(loss)))
        return loss.value


class FoldIn:
    """Fold-in of CSR rows against fixed item factors H [k][n]."""

    def __init__(self, H, lam: float, device: int = 0):
        self._lib = _lib()
        H = _f32(H)
        self.k, self.n = H.shape
        self._f = C.c_void_p()
        rc = self._lib.ccd_fold_create(device, self.k, self.n, N.ptr(H, C.c_float), lam, C.byref(self._f))
        if rc != 0:
            raise CCDError(rc, "ccd_fold_create failed")

    def _check(self, rc):
        if rc < 0:
            raise CCDError(rc, self._lib.ccd_fold_last_error(self._f).decode())

    def close(self):
        if self._f:
            self._lib.ccd_fold_destroy(self._f)
            self._f = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def features(self, ptr, idx, val) -> np.ndarray:
        ptr, idx, val = _u32(ptr), _u32(idx), _f32(val)
        F = np.empty((ptr.size - 1, 2 * self.k + 2), np.float32)
        self._check(self._lib.ccd_fold_features(self._f, ptr.size - 1, N.ptr(ptr, C.c_uint32), N.ptr(idx, C.c_uint32),
                                                N.ptr(val, C.c_float), N.ptr(F, C.c_float)))
        return F

    def rows(self, ptr, idx, val, sweeps: int, W0=None):
        ptr, idx, val = _u32(ptr), _u32(idx), _f32(val)
        W = np.empty((ptr.size - 1, self.k), np.float32)
        keep = _f32(W0) if W0 is not None else None
        w0 = N.ptr(keep, C.c_float) if keep is not None else None
        s = C.c_double()
        self._check(self._lib.ccd_fold_rows(self._f, ptr.size - 1, N.ptr(ptr, C.c_uint32), N.ptr(idx, C.c_uint32),
                                            N.ptr(val, C.c_float), w0, sweeps, N.ptr(W, C.c_float), C.byref(s)))
        return W, s.value

    def rows_noalloc(self, ptr, idx, val, sweeps: int, W0=None):
        """rows() through ccd_fold_rows_noalloc: no device allocation, at most
        4096 rows and 2,097,152 ratings per call."""
        ptr, idx, val = _u32(ptr), _u32(idx), _f32(val)
        W = np.empty((ptr.size - 1, self.k), np.float32)
        keep = _f32(W0) if W0 is not None else None
        w0 = N.ptr(keep, C.c_float) if keep is not None else None
        s = C.c_double()
        self._check(self._lib.ccd_fold_rows_noalloc(self._f, ptr.size - 1, N.ptr(ptr, C.c_uint32),
                                                    N.ptr(idx, C.c_uint32), N.ptr(val, C.c_float), w0, sweeps,
                                                    N.ptr(W, C.c_float), C.byref(s)))
        return W, s.value

    def warm(self, mlp: MLP, ptr, idx, val, sweeps: int, bf16: int = 1):
        """Warm start then `sweeps` sweeps; bf16 is the MLP inference mode
        (0 FP32, 1 bfloat16 weights, 2 matrix cores; ccd_mlp_forward)."""
        ptr, idx, val = _u32(ptr), _u32(idx), _f32(val)
        W = np.empty((ptr.size - 1, self.k), np.float32)
        s = C.c_double()
        self._check(self._lib.ccd_fold_warm(self._f, mlp._m, ptr.size - 1, N.ptr(ptr, C.c_uint32),
                                            N.ptr(idx, C.c_uint32), N.ptr(val, C.c_float), sweeps, int(bf16),
                                            N.ptr(W, C.c_float), C.byref(s)))
        return W, s.value


# ------------------------------------------------------------ experiment ---
@dataclass
class UserSplit:
    old: Dataset          # ratings of the kept users (rows renumbered), test set included
    new_ptr: np.ndarray   # CSR of the held-out users' training ratings
    new_idx: np.ndarray
    new_val: np.ndarray
    new_test_row: np.ndarray  # held-out users' test ratings (row = index among held-out users)
    new_test_col: np.ndarray
    new_test_val: np.ndarray


def split_users(ds: Dataset, frac: float = 0.1, seed: int = 5) -> UserSplit:
    from .datadir import from_triplets
    rng = np.random.default_rng(seed)
    is_new = np.zeros(ds.m, bool)
    is_new[rng.choice(ds.m, size=int(round(frac * ds.m)), replace=False)] = True
    rows = np.repeat(np.arange(ds.m, dtype=np.int64), np.diff(ds.csr_ptr.astype(np.int64)))
    old_id = np.cumsum(~is_new) - 1
    new_id = np.cumsum(is_new) - 1
    keep = ~is_new[rows]
    tkeep = ~is_new[ds.test_row]
    old = from_triplets(int((~is_new).sum()), ds.n,
                        (old_id[rows[keep]], ds.csr_idx[keep], ds.csr_val[keep]),
                        (old_id[ds.test_row[tkeep]], ds.test_col[tkeep], ds.test_val[tkeep]))
    nsel = is_new[rows]
    nrows = int(is_new.sum())
    nr = new_id[rows[nsel]]
    counts = np.bincount(nr, minlength=nrows)
    ptr = np.concatenate([[0], np.cumsum(counts)]).astype(np.uint32)
    return UserSplit(old, ptr, ds.csr_idx[nsel].astype(np.uint32), ds.csr_val[nsel].astype(np.float32),
                     new_id[ds.test_row[~tkeep]].astype(np.uint32), ds.test_col[~tkeep].astype(np.uint32),
                     ds.test_val[~tkeep].astype(np.float32))


def _pred(W, H, rows, cols, chunk=1 << 20):
    out = np.empty(rows.size)
    for a in range(0, rows.size, chunk):
        b = min(rows.size, a + chunk)
        out[a:b] = np.einsum("ij,ji->i", W[rows[a:b]].astype(np.float64), H[:, cols[a:b]].astype(np.float64))
    return out


def fold_objective(W, H, ptr, idx, val, lam) -> float:
    """The fold-in objective sum_i [ sum_j (r_ij - w_i.h_j)^2 + lambda |O_i| ||w_i||^2 ] (FP64)."""
    rows = np.repeat(np.arange(ptr.size - 1), np.diff(ptr.astype(np.int64)))
    d = val.astype(np.float64) - _pred(W, H, rows, idx)
    deg = np.diff(ptr.astype(np.int64)).astype(np.float64)
    return float(np.dot(d, d) + lam * np.dot(deg, (W.astype(np.float64) ** 2).sum(1)))


def fold_rmse(W, H, row, col, val) -> float:
    """Test RMSE of folded-in users: W [users][k], H [k][n] (FP64 sum)."""
    d = _pred(W, H, row, col) - val.astype(np.float64)
    return math.sqrt(float(np.mean(d * d))) if d.size else 0.0


@dataclass
class FoldinResult:
    plain_rmse: list = field(default_factory=list)   # index s: after s sweeps from zero (s = 0 .. S)
    plain_time: list = field(default_factory=list)
    warm_rmse: list = field(default_factory=list)    # index s: after the warm start and s sweeps
    warm_time: list = field(default_factory=list)
    plain_obj: list = field(default_factory=list)    # fold-in objective after s sweeps
    warm_obj: list = field(default_factory=list)
    obj_star: float = float("nan")                   # objective after `converge` plain sweeps
    target: float = float("nan")
    plain_sweeps: int = 0
    warm_sweeps: int = -1
    plain_sweeps_obj: int = -1   # fewest sweeps to within obj_rtol of obj_star
    warm_sweeps_obj: int = -1
    train_loss: float = float("nan")
    timings: dict = field(default_factory=dict)
    # converged target: the test RMSE at the fold-in's fixed point (after
    # `converge` sweeps; both starts reach the same one), and the fewest
    # sweeps after which each start stays within rmse_tol of it
    rmse_star: float = float("nan")
    rmse_star_warm: float = float("nan")
    rmse_tol: float = float("nan")
    plain_sweeps_conv: int = -1
    warm_sweeps_conv: int = -1
    # the same for several tolerances: {tol: (plain sweeps, warm sweeps)}
    conv_by_tol: dict = field(default_factory=dict)


def _sweeps_to_stay_within(curve, star, tol) -> int:
    """Fewest s with |curve[s'] - star| <= tol for every s' >= s (-1: never)."""
    out = -1
    for s in range(len(curve) - 1, -1, -1):
        if abs(curve[s] - star) <= tol:
            out = s
        else:
            break
    return out


def experiment(ds: Dataset, k: int, lam: float, outer: int, inner: int, sweeps: int = 8, epochs: int = 20,
               batch: int = 1024, lr: float = 1e-3, hidden: int = 256, seed: int = 5, converge: int = 40,
               obj_rtol: float = 1e-3, rmse_tol: float = 1e-4, max_sweeps: int = 0, mlp_mode: int = 2,
               log=None) -> FoldinResult:
    """Held-out-users study: factorize the kept users, train the MLP on
    their (features, w_i) pairs, fold in the held-out users from zero and
    from the warm start.  Three targets: the plain fold-in's test RMSE after
    `sweeps` sweeps (the warm start's count is the fewest sweeps that reach
    it); the fold-in objective within obj_rtol of its value after `converge`
    plain sweeps; and the converged test RMSE (after `converge` sweeps from
    either start: per-row fold-in is a convex ridge problem, so both reach
    the same fixed point), for which each start's count is the fewest sweeps
    after which it stays within rmse_tol of it.  Curves are recorded up to
    max(sweeps, max_sweeps) sweeps.  mlp_mode: MLP inference of the warm
    start (1: bfloat16 weights, FP32 arithmetic; 2: the matrix-core path,
    see ccd_mlp_forward)."""
    say = log or (lambda *_: None)
    t = time.perf_counter()
    sp = split_users(ds, 0.1, seed)
    res = FoldinResult()
    res.timings["split_s"] = time.perf_counter() - t
    from .data import initial_col
    with Solver(0) as s:
        s.set_plan(sigma=16)
        s.upload(sp.old)
        s.set_factors(initial_col(k, sp.old.m), np.zeros((k, sp.old.n), np.float32))
        t = time.perf_counter()
        r = s.run(k, lam, outer, inner, variant=VARIANT_TILED)
        res.timings["solve_s"] = time.perf_counter() - t
        W, H = s.factors()
    say("solved %d kept users: test RMSE %.5f" % (sp.old.m, r.rmse[-1] if r.rmse else float("nan")))
    fold = FoldIn(H, lam)
    X = fold.features(sp.old.csr_ptr, sp.old.csr_idx, sp.old.csr_val)
    T = np.ascontiguousarray(W.T)
    mlp = MLP(2 * k + 2, hidden, k, seed=seed)
    mlp.set_norm(X.mean(0), X.std(0) + 1e-6, T.mean(0), T.std(0) + 1e-6)
    t = time.perf_counter()
    res.train_loss = mlp.train(X, T, epochs=epochs, batch=batch, lr=lr, seed=seed)
    res.timings["train_s"] = time.perf_counter() - t
    say("MLP trained on %d users, %d epochs: loss %.4f (%.1f s)" % (X.shape[0], epochs, res.train_loss,
                                                                     res.timings["train_s"]))
    args = (sp.new_ptr, sp.new_idx, sp.new_val)
    test = (sp.new_test_row, sp.new_test_col, sp.new_test_val)
    fold.warm(mlp, *args, sweeps=0, bf16=mlp_mode)  # first call: buffers, bf16 weights
    Wc, _ = fold.rows(*args, sweeps=converge)
    res.obj_star = fold_objective(Wc, H, *args, lam)
    res.rmse_star = fold_rmse(Wc, H, *test)
    Wcw, _ = fold.warm(mlp, *args, sweeps=converge, bf16=mlp_mode)
    res.rmse_star_warm = fold_rmse(Wcw, H, *test)
    res.rmse_tol = rmse_tol
    say("fixed point after %d sweeps: test RMSE %.6f from zero, %.6f from the warm start" % (
        converge, res.rmse_star, res.rmse_star_warm))
    for sw in range(max(sweeps, max_sweeps) + 1):
        Wp, tp = fold.rows(*args, sweeps=sw)
        res.plain_rmse.append(fold_rmse(Wp, H, *test))
        res.plain_time.append(tp)
        Ww, tw = fold.warm(mlp, *args, sweeps=sw, bf16=mlp_mode)
        res.warm_rmse.append(fold_rmse(Ww, H, *test))
        res.warm_time.append(tw)
        if sw <= sweeps:  # the objective (a pass over every rating on the host) for the first sweeps only
            res.plain_obj.append(fold_objective(Wp, H, *args, lam))
            res.warm_obj.append(fold_objective(Ww, H, *args, lam))
            say("sweeps %d: plain %.5f obj %.6g (%.2f ms)  warm %.5f obj %.6g (%.2f ms)" % (
                sw, res.plain_rmse[-1], res.plain_obj[-1] / res.obj_star - 1, 1e3 * tp, res.warm_rmse[-1],
                res.warm_obj[-1] / res.obj_star - 1, 1e3 * tw))
        else:
            say("sweeps %d: plain %.6f (%.2f ms)  warm %.6f (%.2f ms)" % (
                sw, res.plain_rmse[-1], 1e3 * tp, res.warm_rmse[-1], 1e3 * tw))
    res.plain_sweeps_conv = _sweeps_to_stay_within(res.plain_rmse, res.rmse_star, rmse_tol)
    res.warm_sweeps_conv = _sweeps_to_stay_within(res.warm_rmse, res.rmse_star, rmse_tol)
    for tol in sorted({1e-3, 5e-4, 2e-4, 1e-4, rmse_tol}, reverse=True):
        res.conv_by_tol[tol] = (_sweeps_to_stay_within(res.plain_rmse, res.rmse_star, tol),
                                _sweeps_to_stay_within(res.warm_rmse, res.rmse_star, tol))
        say("within %.0e of the converged test RMSE from: plain %d sweeps, warm %d sweeps" % (
            tol, *res.conv_by_tol[tol]))
    say("within %.0e of the converged test RMSE from: plain %d sweeps, warm %d sweeps" % (
        rmse_tol, res.plain_sweeps_conv, res.warm_sweeps_conv))
    res.target = res.plain_rmse[sweeps]
    res.plain_sweeps = sweeps
    res.warm_sweeps = next((sw for sw, v in enumerate(res.warm_rmse) if v <= res.target), -1)
    res.plain_sweeps_obj = next((sw for sw, v in enumerate(res.plain_obj) if v <= res.obj_star * (1 + obj_rtol)), -1)
    res.warm_sweeps_obj = next((sw for sw, v in enumerate(res.warm_obj) if v <= res.obj_star * (1 + obj_rtol)), -1)
    fold.close()
    mlp.close()
    return res
