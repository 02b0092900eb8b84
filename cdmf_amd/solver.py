"""Python side of the C ABI in include/cdmf_hip.h.

`Solver` wraps one device context.  `cdmf_hip` and the reference-named
`cdmf_ocl_01`, `cdmf_ocl_02`, `cdmf_csr5` restore the reference's one-call
solver signature (src/cdmf.h:4-10): they take the matrix, rank-major W and H
(numpy [k][m] / [k][n], the MatData layout of src/pmf_util.h:7-8), the test
set and a `Parameter` (src/pmf.h:4-26); they zero H, run the solve on the
GPU and write W and H back in place.

There is no CPU path here: without the built library or without a GPU these
raise.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field

import numpy as np

from . import _native as N
from .datadir import Dataset

# enum values of include/cdmf_hip.h
VARIANT_NATIVE, VARIANT_BATCHING, VARIANT_TILED = 1, 2, 3
PREC_FP32, PREC_BF16, PREC_FP64_SUMS = 0, 1, 2
ORDER_H_FIRST, ORDER_W_FIRST = 0, 1
K_SWEEP_F, K_SWEEP_S, K_FUSED_F, K_FUSED_S, K_RESIDUAL, K_FINALIZE, K_RMSE, K_COPY = range(8)
KERNEL_KINDS = ("sweep_f", "sweep_s", "fused_f", "fused_s", "residual", "finalize", "rmse", "copy")


class CParams(C.Structure):
    _fields_ = [
        ("k", C.c_uint32), ("lambda_", C.c_float), ("maxiter", C.c_int32), ("maxinneriter", C.c_int32),
        ("variant", C.c_int32), ("precision", C.c_int32), ("order", C.c_int32), ("zero_first", C.c_int32),
        ("rmse_every", C.c_int32), ("profile", C.c_int32), ("use_graph", C.c_int32),
        ("profile_kinds", C.c_int32), ("reserved", C.c_int32 * 4),
    ]


class CPlan(C.Structure):
    _fields_ = [("elems_per_lane", C.c_uint32), ("split_threshold", C.c_uint32), ("piece", C.c_uint32),
                ("unit_elems", C.c_uint32), ("sigma", C.c_uint32), ("blocks_per_cu", C.c_uint32),
                ("split_rows", C.c_uint32)]


class CBatchPlan(C.Structure):
    _fields_ = [("sigma", C.c_uint32), ("blocks", C.c_uint32), ("work_exponent", C.c_float), ("solo", C.c_uint32),
                ("deal", C.c_uint32)]


ITER_CB = C.CFUNCTYPE(None, C.c_void_p, C.c_int32, C.c_double, C.c_double, C.c_double, C.c_double)

_hip = None


ABI_VERSION = 6  # CCD_ABI_VERSION of include/cdmf_hip.h this binding is written against


def hip_lib() -> C.CDLL:
    global _hip
    if _hip is None:
        lib = N._open("libcdmf_hip.so")
        vp = C.c_void_p
        u32p, f32p, f64p = N.u32p, N.f32p, N.f64p
        sig = {
            "ccd_abi_version": ([], C.c_int32),
            "ccd_device_count": ([], C.c_int32),
            "ccd_params_default": ([C.POINTER(CParams)], None),
            "ccd_create": ([C.c_int32, C.POINTER(vp)], C.c_int32),
            "ccd_destroy": ([vp], C.c_int32),
            "ccd_last_error": ([vp], C.c_char_p),
            "ccd_upload": ([vp, C.c_int64, C.c_int64, C.c_uint64, u32p, u32p, f32p, u32p, u32p, f32p], C.c_int32),
            "ccd_upload_test": ([vp, C.c_uint64, u32p, u32p, f32p], C.c_int32),
            "ccd_set_factors": ([vp, C.c_uint32, f32p, f32p], C.c_int32),
            "ccd_run": ([vp, C.POINTER(CParams), ITER_CB, vp], C.c_int32),
            "ccd_get_factors": ([vp, f32p, f32p], C.c_int32),
            "ccd_get_residual": ([vp, f32p, f32p], C.c_int32),
            "ccd_rmse": ([vp, f64p], C.c_int32),
            "ccd_objective": ([vp, C.c_float, f64p], C.c_int32),
            "ccd_iter_rmse": ([vp, f64p, C.c_int32], C.c_int32),
            "ccd_kernel_stats": ([vp, C.c_int32, f64p, C.POINTER(C.c_uint64), f64p], C.c_int32),
            "ccd_loop_time": ([vp, f64p], C.c_int32),
            "ccd_copy_probe": ([vp, C.c_uint64, C.c_int32, f64p], C.c_int32),
            "ccd_set_plan": ([vp, C.POINTER(CPlan)], C.c_int32),
            "ccd_batch_create": ([C.c_int32, C.POINTER(vp)], C.c_int32),
            "ccd_batch_destroy": ([vp], C.c_int32),
            "ccd_batch_last_error": ([vp], C.c_char_p),
            "ccd_batch_add": ([vp, vp, C.POINTER(CParams)], C.c_int32),
            "ccd_batch_set_plan": ([vp, C.POINTER(CBatchPlan)], C.c_int32),
            "ccd_batch_get_plan": ([vp, C.POINTER(CBatchPlan)], C.c_int32),
            "ccd_batch_tune": ([vp, f64p], C.c_int32),
            "ccd_batch_run": ([vp, C.c_int32], C.c_int32),
            "ccd_batch_time": ([vp, f64p], C.c_int32),
            "ccd_batch_teams": ([vp, u32p, C.c_int32], C.c_int32),
            "ccd_batch_tune_times": ([vp, f64p, C.c_int32], C.c_int32),
        }
        for name, (args, res) in sig.items():
            f = getattr(lib, name)
            f.argtypes = args
            f.restype = res
        got = lib.ccd_abi_version()
        if got != ABI_VERSION:
            raise RuntimeError("libcdmf_hip.so has ABI version %d, this binding expects %d (rebuild the library)"
                               % (got, ABI_VERSION))
        _hip = lib
    return _hip


# Every symbol declared in include/cdmf_hip.h.
ABI_SYMBOLS = (
    "ccd_abi_version", "ccd_device_count", "ccd_params_default", "ccd_create", "ccd_destroy", "ccd_last_error",
    "ccd_upload", "ccd_upload_test", "ccd_set_factors", "ccd_run", "ccd_get_factors", "ccd_get_residual",
    "ccd_rmse", "ccd_objective", "ccd_iter_rmse", "ccd_kernel_stats", "ccd_loop_time", "ccd_copy_probe", "ccd_set_plan",
    "ccd_batch_create", "ccd_batch_destroy", "ccd_batch_last_error", "ccd_batch_add", "ccd_batch_set_plan",
    "ccd_batch_get_plan", "ccd_batch_tune", "ccd_batch_run", "ccd_batch_time", "ccd_batch_teams",
    "ccd_batch_tune_times",
)


class CCDError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__("ccd error %d: %s" % (code, msg))
        self.code = code


@dataclass
class Parameter:
    """The reference's solver parameters (src/pmf.h:4-26) plus the options of
    this implementation.  `maxiter` is the outer count (-t), `maxinneriter`
    the inner count (-T)."""

    threads: int = 16
    nBlocks: int = 1
    nThreadsPerBlock: int = 32
    do_ref: int = 0
    do_predict: int = 0
    maxinneriter: int = 5
    maxiter: int = 5
    k: int = 10
    lambda_: float = 0.1
    platform_id: int = 0
    verbose: int = 0
    version: int = 1
    # extensions
    precision: int = PREC_FP32
    order: int = ORDER_H_FIRST
    zero_first: int = 1
    profile: int = 1
    use_graph: int = 0


@dataclass
class IterRecord:
    oiter: int
    rank_time: float
    update_time: float
    rmse: float
    rmse_time: float


@dataclass
class RunResult:
    iters: list = field(default_factory=list)
    rmse: list = field(default_factory=list)


def device_count() -> int:
    return int(hip_lib().ccd_device_count())


class Solver:
    """One device context: R uploaded once, factors and runs on top."""

    def __init__(self, device: int = 0):
        self._lib = hip_lib()
        self._ctx = C.c_void_p()
        rc = self._lib.ccd_create(device, C.byref(self._ctx))
        if rc != 0:
            raise CCDError(rc, "ccd_create(%d) failed (no usable HIP device?)" % device)
        self.m = self.n = self.nnz = 0
        self.k = 0

    def _check(self, rc: int) -> None:
        if rc < 0:
            raise CCDError(rc, self._lib.ccd_last_error(self._ctx).decode())

    def close(self) -> None:
        if self._ctx:
            self._lib.ccd_destroy(self._ctx)
            self._ctx = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_plan(self, elems_per_lane=4, split_threshold=8192, piece=4096, unit_elems=4096, sigma=16,
                 blocks_per_cu=0, split_rows=1) -> None:
        pp = CPlan(elems_per_lane, split_threshold, piece, unit_elems, sigma, blocks_per_cu, split_rows)
        self._check(self._lib.ccd_set_plan(self._ctx, C.byref(pp)))

    def upload(self, ds: Dataset, with_test: bool = True) -> None:
        a = [np.ascontiguousarray(x) for x in (ds.csr_ptr, ds.csr_idx, ds.csr_val, ds.csc_ptr, ds.csc_idx, ds.csc_val)]
        self._check(self._lib.ccd_upload(
            self._ctx, ds.m, ds.n, ds.nnz,
            N.ptr(a[0].astype(np.uint32, copy=False), C.c_uint32), N.ptr(a[1].astype(np.uint32, copy=False), C.c_uint32),
            N.ptr(a[2].astype(np.float32, copy=False), C.c_float), N.ptr(a[3].astype(np.uint32, copy=False), C.c_uint32),
            N.ptr(a[4].astype(np.uint32, copy=False), C.c_uint32), N.ptr(a[5].astype(np.float32, copy=False), C.c_float)))
        self.m, self.n, self.nnz = ds.m, ds.n, ds.nnz
        if with_test:
            self.upload_test(ds.test_row, ds.test_col, ds.test_val)

    def upload_test(self, row, col, val) -> None:
        r = np.ascontiguousarray(row, dtype=np.uint32)
        c = np.ascontiguousarray(col, dtype=np.uint32)
        v = np.ascontiguousarray(val, dtype=np.float32)
        self._check(self._lib.ccd_upload_test(self._ctx, r.size, N.ptr(r, C.c_uint32), N.ptr(c, C.c_uint32),
                                              N.ptr(v, C.c_float)))

    def set_factors(self, W: np.ndarray, H: np.ndarray) -> None:
        W = np.ascontiguousarray(W, dtype=np.float32)
        H = np.ascontiguousarray(H, dtype=np.float32)
        k = W.shape[0]
        if W.shape != (k, self.m) or H.shape != (k, self.n):
            raise ValueError("W must be [k][m] and H [k][n]")
        self._check(self._lib.ccd_set_factors(self._ctx, k, N.ptr(W, C.c_float), N.ptr(H, C.c_float)))
        self.k = k

    def run(self, k: int, lam: float, outer: int, inner: int, variant: int = VARIANT_NATIVE,
            precision: int = PREC_FP32, order: int = ORDER_H_FIRST, zero_first: bool = True,
            rmse_every: bool = True, profile: bool = False, use_graph: bool = False,
            callback=None, profile_kinds=()) -> RunResult:
        """profile_kinds: names from KERNEL_KINDS to time when profile is set
        (empty: every kind)."""
        p = CParams()
        self._lib.ccd_params_default(C.byref(p))
        p.k, p.lambda_, p.maxiter, p.maxinneriter = k, lam, outer, inner
        p.variant, p.precision, p.order = variant, precision, order
        p.zero_first, p.rmse_every, p.profile, p.use_graph = int(zero_first), int(rmse_every), int(profile), int(use_graph)
        p.profile_kinds = sum(1 << KERNEL_KINDS.index(kk) for kk in set(profile_kinds))
        res = RunResult()

        def _cb(_user, oiter, rank_t, upd_t, rmse, rmse_t):
            rec = IterRecord(oiter, rank_t, upd_t, rmse, rmse_t)
            res.iters.append(rec)
            if callback is not None:
                callback(rec)

        cb = ITER_CB(_cb) if callback is not None else ITER_CB()
        self._check(self._lib.ccd_run(self._ctx, C.byref(p), cb, None))
        buf = np.full(max(outer, 1), np.nan)
        n = self._lib.ccd_iter_rmse(self._ctx, N.ptr(buf, C.c_double), outer)
        res.rmse = [float(x) for x in buf[:max(n, 0)]]
        return res

    def factors(self):
        W = np.empty((self.k, self.m), dtype=np.float32)
        H = np.empty((self.k, self.n), dtype=np.float32)
        self._check(self._lib.ccd_get_factors(self._ctx, N.ptr(W, C.c_float), N.ptr(H, C.c_float)))
        return W, H

    def residual(self):
        a = np.empty(self.nnz, dtype=np.float32)
        b = np.empty(self.nnz, dtype=np.float32)
        self._check(self._lib.ccd_get_residual(self._ctx, N.ptr(a, C.c_float), N.ptr(b, C.c_float)))
        return a, b

    def rmse(self) -> float:
        out = C.c_double()
        self._check(self._lib.ccd_rmse(self._ctx, C.byref(out)))
        return out.value

    def objective(self, lam: float):
        """(objective, loss, regulariser) of the current state, on the device
        (include/cdmf_hip.h, ccd_objective)."""
        out = np.zeros(3, dtype=np.float64)
        self._check(self._lib.ccd_objective(self._ctx, lam, N.ptr(out, C.c_double)))
        return out[0], out[1], out[2]

    def kernel_stats(self) -> dict:
        out = {}
        for kind, name in enumerate(KERNEL_KINDS):
            ms, n, b = C.c_double(), C.c_uint64(), C.c_double()
            self._check(self._lib.ccd_kernel_stats(self._ctx, kind, C.byref(ms), C.byref(n), C.byref(b)))
            out[name] = dict(total_ms=ms.value, launches=n.value, bytes_per_launch=b.value)
        return out

    def loop_time(self) -> float:
        s = C.c_double()
        self._check(self._lib.ccd_loop_time(self._ctx, C.byref(s)))
        return s.value

    def copy_probe(self, nbytes: int, reps: int = 20) -> float:
        ms = C.c_double()
        self._check(self._lib.ccd_copy_probe(self._ctx, nbytes, reps, C.byref(ms)))
        return ms.value


def make_params(k: int, lam: float, outer: int, inner: int, variant: int = VARIANT_TILED,
                precision: int = PREC_FP32, order: int = ORDER_H_FIRST, zero_first: bool = True) -> CParams:
    p = CParams()
    hip_lib().ccd_params_default(C.byref(p))
    p.k, p.lambda_, p.maxiter, p.maxinneriter = k, lam, outer, inner
    p.variant, p.precision, p.order, p.zero_first = variant, precision, order, int(zero_first)
    return p


class Batch:
    """Many independent factorizations solved by one persistent launch
    (include/cdmf_hip.h, ccd_batch_*).  Each problem is a Solver with R
    uploaded and factors set; results are read back from the Solvers."""

    def __init__(self, device: int = 0):
        self._lib = hip_lib()
        self._b = C.c_void_p()
        rc = self._lib.ccd_batch_create(device, C.byref(self._b))
        if rc != 0:
            raise CCDError(rc, "ccd_batch_create(%d) failed" % device)
        self.solvers = []

    def _check(self, rc: int) -> int:
        if rc < 0:
            raise CCDError(rc, self._lib.ccd_batch_last_error(self._b).decode())
        return rc

    def close(self) -> None:
        if self._b:
            self._lib.ccd_batch_destroy(self._b)
            self._b = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add(self, solver: Solver, k: int, lam: float, outer: int, inner: int, order: int = ORDER_H_FIRST,
            zero_first: bool = True) -> int:
        p = make_params(k, lam, outer, inner, VARIANT_TILED, PREC_FP32, order, zero_first)
        idx = self._check(self._lib.ccd_batch_add(self._b, solver._ctx, C.byref(p)))
        self.solvers.append(solver)
        return idx

    def set_plan(self, sigma: int = 16, blocks: int = 0, work_exponent: float = 1.0, solo: int = 0,
                 deal: int = 0) -> None:
        """solo: the largest problems solved alone on the whole GPU first
        (ccd_batch_plan.solo); deal: tiles per wave from which a team's
        sweep deals its tiles through an LDS counter (0: the default, 16;
        0xffffffff: never)."""
        self._check(self._lib.ccd_batch_set_plan(self._b, C.byref(CBatchPlan(sigma, blocks, work_exponent, solo,
                                                                              deal))))

    def plan(self) -> dict:
        pp = CBatchPlan()
        self._check(self._lib.ccd_batch_get_plan(self._b, C.byref(pp)))
        return dict(sigma=pp.sigma, blocks=pp.blocks, work_exponent=pp.work_exponent, solo=pp.solo, deal=pp.deal)

    def tune(self) -> float:
        ms = C.c_double()
        self._check(self._lib.ccd_batch_tune(self._b, C.byref(ms)))
        return ms.value

    def run(self, use_graph: bool = False) -> float:
        """One run of every problem; returns its time in seconds
        (ccd_batch_time)."""
        self._check(self._lib.ccd_batch_run(self._b, int(use_graph)))
        s = C.c_double()
        self._check(self._lib.ccd_batch_time(self._b, C.byref(s)))
        return s.value

    def tune_times(self) -> list:
        """Whole-run times (ms) of the last tune: solo = 0, 1, ..., min(3,
        problems - 1), then deal = 8, 32, never (ccd_batch_tune_times)."""
        buf = np.zeros(16)
        n = self._check(self._lib.ccd_batch_tune_times(self._b, N.ptr(buf, C.c_double), buf.size))
        return [float(x) for x in buf[:min(n, buf.size)]]

    def teams(self) -> list:
        buf = np.zeros(max(len(self.solvers), 1), dtype=np.uint32)
        n = self._check(self._lib.ccd_batch_teams(self._b, N.ptr(buf, C.c_uint32), buf.size))
        return [int(x) for x in buf[:n]]


def cdmf_hip(R: Dataset, W: np.ndarray, H: np.ndarray, T: Dataset | None, param: Parameter,
             device: int = 0, variant: int | None = None, log=None) -> RunResult:
    """One-call solve with the reference's semantics (cdmf_ocl_02, src/cdmf_ocl_2.cpp:6-318):
    H (the first-swept factor) is zeroed, W is used as given, W and H are
    updated in place.  `log(line)` receives the per-iteration line of
    cdmf_ocl_2.cpp:291-292."""
    with Solver(device) as s:
        s.upload(R, with_test=False)
        if T is not None and T.nnz_test:
            s.upload_test(T.test_row, T.test_col, T.test_val)
        s.set_factors(W, H)
        acc = [0.0, 0.0]

        def cb(rec: IterRecord):
            acc[0] += rec.rank_time
            acc[1] += rec.update_time
            if log is not None:
                log("[-INFO-] iteration num %d \trank_time %.4lf|%.4lf s \tupdate_time %.4lf|%.4lfs \tRMSE=%f time:%fs"
                    % (rec.oiter, rec.rank_time, acc[0], rec.update_time, acc[1], rec.rmse, rec.rmse_time))

        res = s.run(param.k, param.lambda_, param.maxiter, param.maxinneriter,
                    variant=param.version if variant is None else variant, precision=param.precision,
                    order=param.order, zero_first=bool(param.zero_first), rmse_every=T is not None,
                    profile=bool(param.profile), use_graph=bool(param.use_graph), callback=cb)
        Wn, Hn = s.factors()
        W[...] = Wn
        H[...] = Hn
        return res


def cdmf_ocl_01(R, W, H, T, param, filename=None, **kw):
    """Drop-in for src/cdmf_ocl_1.cpp:6 (the kernel file name is unused)."""
    return cdmf_hip(R, W, H, T, param, variant=VARIANT_NATIVE, **kw)


def cdmf_ocl_02(R, W, H, T, param, filename=None, **kw):
    """Drop-in for src/cdmf_ocl_2.cpp:6 (the kernel file name is unused)."""
    return cdmf_hip(R, W, H, T, param, variant=VARIANT_BATCHING, **kw)


def cdmf_csr5(R, W, H, param, filename=None, **kw):
    """Drop-in for src/cdmf_csr5.cpp:7: no test set, tiled sweeps."""
    return cdmf_hip(R, W, H, None, param, variant=VARIANT_TILED, **kw)


def golden_compare(W: np.ndarray, W_ref: np.ndarray) -> tuple[int, float]:
    """src/tools.cpp:467-487: count entries with |W - W_ref| > 0.1 |W_ref|;
    returns (error_count, error_percentage)."""
    delta = np.abs(W.astype(np.float64) - W_ref.astype(np.float64))
    err = int(np.count_nonzero(delta > 0.1 * np.abs(W_ref.astype(np.float64))))
    return err, 100.0 * err / max(W.size, 1)


def rmse_host(W: np.ndarray, H: np.ndarray, row, col, val) -> float:
    """Test RMSE from downloaded factors (tools.cpp:543-582 semantics) in numpy."""
    pred = np.zeros(len(row), dtype=np.float64)
    for t in range(W.shape[0]):
        pred += (W[t, row] * H[t, col]).astype(np.float32).astype(np.float64)
    d = (pred - val.astype(np.float64)) ** 2
    d = np.where(np.isnan(d), 0.0, d)
    return math.sqrt(d.sum() / max(len(row), 1))
