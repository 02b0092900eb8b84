"""CPU pins of the fold-in / MLP oracle (oracle/foldin_oracle.c).

The reference has no fold-in, so the oracle is pinned by mathematics:
  * CCD++ fold-in sweeps converge to the ridge solve
    (H_O^T H_O + lambda |O| I) w = H_O^T r of every row (numpy solve);
  * a warm start at that solution is a fixed point of a sweep;
  * the features equal a numpy computation that uses E[h^2] - E[h]^2 for the
    variance (different arithmetic from the oracle's two-pass form);
  * the MLP forward equals a numpy/scipy forward, and its gradient equals
    central finite differences of the loss.
"""
import numpy as np
import pytest
from scipy.special import erf


def _rows(seed=0, rows=40, n=60, k=5):
    rng = np.random.default_rng(seed)
    deg = rng.integers(0, 25, size=rows)
    deg[3] = 0  # an empty row
    ptr = np.concatenate([[0], np.cumsum(deg)]).astype(np.uint32)
    idx = np.concatenate([np.sort(rng.choice(n, size=d, replace=False)) for d in deg]).astype(np.uint32)
    val = rng.integers(1, 6, size=idx.size).astype(np.float32)
    H = (rng.random((k, n)) * 0.6).astype(np.float32)
    return ptr, idx, val, H


def _ridge(ptr, idx, val, H, lam):
    k = H.shape[0]
    out = np.zeros((ptr.size - 1, k))
    for i in range(ptr.size - 1):
        s, e = int(ptr[i]), int(ptr[i + 1])
        if e == s:
            continue
        Ho = H[:, idx[s:e]].astype(np.float64)
        A = Ho @ Ho.T + lam * (e - s) * np.eye(k)
        out[i] = np.linalg.solve(A, Ho @ val[s:e].astype(np.float64))
    return out


def test_fold_rows_reach_ridge_solution(oracle_lib):
    ptr, idx, val, H = _rows()
    lam = 0.05
    W = oracle_lib.fold_rows(ptr, idx, val, H, lam, sweeps=400)
    ref = _ridge(ptr, idx, val, H, lam)
    np.testing.assert_allclose(W, ref, rtol=2e-4, atol=2e-4)
    assert np.all(W[3] == 0.0)


def test_fold_rows_fixed_point_and_warm_start(oracle_lib):
    ptr, idx, val, H = _rows(seed=1)
    lam = 0.1
    ref = _ridge(ptr, idx, val, H, lam).astype(np.float32)
    W1 = oracle_lib.fold_rows(ptr, idx, val, H, lam, sweeps=1, W0=ref)
    np.testing.assert_allclose(W1, ref, rtol=1e-4, atol=1e-5)
    # one sweep from the solution beats one sweep from zero by a wide margin
    Wz = oracle_lib.fold_rows(ptr, idx, val, H, lam, sweeps=1)
    assert np.abs(W1 - ref).max() < 0.1 * np.abs(Wz - ref).max()


def test_fold_features_match_numpy(oracle_lib):
    ptr, idx, val, H = _rows(seed=2)
    k = H.shape[0]
    F = oracle_lib.fold_features(ptr, idx, val, H)
    for i in range(ptr.size - 1):
        s, e = int(ptr[i]), int(ptr[i + 1])
        r = val[s:e].astype(np.float64)
        Ho = H[:, idx[s:e]].astype(np.float64)
        if e > s:
            m = (Ho * r).sum(1) / r.sum()
            v = (Ho * Ho * r).sum(1) / r.sum() - m * m
            np.testing.assert_allclose(F[i, :k], m, rtol=1e-5, atol=1e-7)
            np.testing.assert_allclose(F[i, k:2 * k], v, rtol=1e-4, atol=1e-6)
            assert F[i, 2 * k + 1] == pytest.approx(r.mean(), rel=1e-6)
        else:
            assert np.all(F[i, :2 * k + 2] == 0.0)
        assert F[i, 2 * k] == pytest.approx(np.log1p(e - s), rel=1e-6)


def _np_forward(P, X, dims):
    nin, hid, out = dims
    o = 0
    W1 = P[o:o + hid * nin].reshape(hid, nin); o += hid * nin
    b1 = P[o:o + hid]; o += hid
    W2 = P[o:o + hid * hid].reshape(hid, hid); o += hid * hid
    b2 = P[o:o + hid]; o += hid
    W3 = P[o:o + out * hid].reshape(out, hid); o += out * hid
    b3 = P[o:o + out]
    g = lambda z: 0.5 * z * (1.0 + erf(z / np.sqrt(2.0)))
    return g(g(X @ W1.T + b1) @ W2.T + b2) @ W3.T + b3


def test_mlp_forward_and_gradient(oracle_lib):
    dims = (6, 5, 3)
    nin, hid, out = dims
    npar = hid * nin + hid + hid * hid + hid + out * hid + out
    rng = np.random.default_rng(3)
    P = rng.normal(0, 0.5, npar)
    X = rng.normal(0, 1, (7, nin))
    T = rng.normal(0, 1, (7, out))
    np.testing.assert_allclose(oracle_lib.mlp_forward(P, X, dims), _np_forward(P, X, dims), rtol=1e-12, atol=1e-12)
    loss, G = oracle_lib.mlp_backward(P, X, T, dims)
    assert loss == pytest.approx(np.mean((_np_forward(P, X, dims) - T) ** 2), rel=1e-12)
    eps = 1e-6
    num = np.empty(npar)
    for q in range(npar):
        Pp, Pm = P.copy(), P.copy()
        Pp[q] += eps
        Pm[q] -= eps
        num[q] = (np.mean((_np_forward(Pp, X, dims) - T) ** 2) - np.mean((_np_forward(Pm, X, dims) - T) ** 2)) / (2 * eps)
    np.testing.assert_allclose(G, num, rtol=1e-6, atol=1e-9)
