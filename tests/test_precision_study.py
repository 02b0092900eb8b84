"""Precision study on the 138,493 x 26,744 shape (18.0M training / 2.0M test
half-star ratings, power-law degrees), k = 64, lambda = 0.05, 10 outer x 5
inner: the GPU solve with FP32 factors, with BF16-stored factors (FP32
arithmetic) and with FP64 per-row sums, each against the FP64 oracle run on
the same seeded matrix and initial factors.  Prints the test RMSE of every
outer iteration and its error against FP64.

Bars: FP32 and FP64-sums within 1e-3 of the FP64 oracle's test RMSE at every
outer iteration (the FP32 target of the large runs); BF16 storage within
1e-2 (bfloat16 keeps 8 significand bits: each stored factor is off by up to
2^-9 relative).
"""
import json
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_precision_study_ml20m_shape(oracle_lib):
    from cdmf_amd import data as D
    from cdmf_amd import solver as S
    if S.device_count() < 1:
        pytest.fail("no HIP device visible: GPU tests need the MI355X")
    ds = D.ml20m()
    cfg = D.ML20M_SOLVE
    k, lam, T, inner = cfg.k, float(np.float32(cfg.lam)), cfg.outer, cfg.inner
    W0 = D.initial_col(k, ds.m)
    H0 = np.zeros((k, ds.n), np.float32)
    out = {}
    with S.Solver(0) as s:
        s.set_plan(sigma=16)
        s.upload(ds)
        for name, prec in (("fp32", S.PREC_FP32), ("bf16", S.PREC_BF16), ("fp64_sums", S.PREC_FP64_SUMS)):
            s.set_factors(W0, H0)
            t = time.perf_counter()
            r = s.run(k, lam, T, inner, variant=S.VARIANT_TILED, precision=prec)
            out[name] = dict(rmse=r.rmse, seconds=time.perf_counter() - t)
            print("%-10s %s (%.2f s)" % (name, ["%.5f" % x for x in r.rmse], out[name]["seconds"]), flush=True)
    t = time.perf_counter()
    _, _, rm64, _, _ = oracle_lib.ccd(ds, W0, H0, k, lam, T, inner, dtype="f64", threads=16)
    out["fp64_oracle"] = dict(rmse=[float(x) for x in rm64], seconds=time.perf_counter() - t)
    print("%-10s %s (%.1f s, CPU)" % ("fp64", ["%.5f" % x for x in rm64], out["fp64_oracle"]["seconds"]), flush=True)
    err = {}
    for name in ("fp32", "bf16", "fp64_sums"):
        err[name] = [abs(a - b) for a, b in zip(out[name]["rmse"], rm64)]
        print("error %-10s %s" % (name, ["%.1e" % e for e in err[name]]), flush=True)
    print("PRECISION_STUDY_JSON " + json.dumps(dict(config=dict(m=ds.m, n=ds.n, nnz=ds.nnz, nnz_test=ds.nnz_test, k=k,
                                                                   lam=cfg.lam, outer=T, inner=inner),
                                                     runs=out, error_vs_fp64=err)), flush=True)
    assert max(err["fp32"]) <= 1e-3
    assert max(err["fp64_sums"]) <= 1e-3
    assert max(err["bf16"]) <= 1e-2
