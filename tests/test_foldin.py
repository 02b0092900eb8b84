"""GPU checks of the fold-in extension against the CPU oracle
(oracle/foldin_oracle.c, pinned in tests/test_oracle_foldin.py).

Bars:
  * features: FP64 sums in the same element order as the oracle, so
    relative 1e-6 (the float rounding of the same doubles);
  * fold-in sweeps: FP32 with a different summation order (wave
    reduction): Frobenius relative 1e-4 and elementwise 1e-3 relative +
    1e-4 rms, from zero and from a warm start, short rows (registers) and
    rows above 1024 ratings (streamed residual); 300 sweeps reach the ridge
    solve within 1e-3;
  * MLP: FP32 GPU forward and gradients against the FP64 oracle at the
    production shape (82 -> 256 -> 256 -> 40), on 96 rows and on 1024 rows
    (the weight-gradient products split K): relative Frobenius 1e-4 per
    parameter tensor, loss within 1e-5 relative; bfloat16-weight inference
    against the oracle run on the bfloat16-rounded weights, 1e-4;
  * training lowers the loss on a smooth regression target tenfold.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _gpu():
    from cdmf_amd import foldin, solver
    if solver.device_count() < 1:
        pytest.fail("no HIP device visible: GPU tests need the MI355X")
    return foldin


def _say(msg):
    print(msg, flush=True)


def _rows(seed, rows, n, k, long_rows=0):
    rng = np.random.default_rng(seed)
    deg = rng.integers(0, 300, size=rows)
    deg[1] = 0
    for q in range(long_rows):
        deg[2 + q] = 1500 + 700 * q  # above 64 * 16: the streamed path
    deg = np.minimum(deg, n)
    ptr = np.concatenate([[0], np.cumsum(deg)]).astype(np.uint32)
    idx = np.concatenate([np.sort(rng.choice(n, size=d, replace=False)) for d in deg]).astype(np.uint32)
    val = rng.integers(1, 6, size=idx.size).astype(np.float32)
    H = (rng.random((k, n)) * 0.5).astype(np.float32)
    return ptr, idx, val, H


def _close(x, ref, what, frob=1e-4, rtol=1e-3):
    x = np.asarray(x, np.float64)
    ref = np.asarray(ref, np.float64)
    rms = float(np.sqrt(np.mean(ref * ref))) if ref.size else 0.0
    f = np.linalg.norm(x - ref) / max(np.linalg.norm(ref), 1e-300)
    assert f <= frob, "%s: Frobenius relative %.3g > %g" % (what, f, frob)
    np.testing.assert_allclose(x, ref, rtol=rtol, atol=1e-4 * rms, err_msg=what)


def test_features_match_oracle(oracle_lib):
    F = _gpu()
    ptr, idx, val, H = _rows(0, 500, 3000, 40, long_rows=2)
    fold = F.FoldIn(H, 0.05)
    got = fold.features(ptr, idx, val)
    ref = oracle_lib.fold_features(ptr, idx, val, H)
    np.testing.assert_allclose(got, ref, rtol=1e-6, atol=1e-7)
    fold.close()


@pytest.mark.parametrize("sweeps", [1, 5])
def test_fold_rows_match_oracle(oracle_lib, sweeps):
    F = _gpu()
    ptr, idx, val, H = _rows(1, 700, 4000, 40, long_rows=2)
    lam = 0.05
    fold = F.FoldIn(H, lam)
    W, t = fold.rows(ptr, idx, val, sweeps)
    _say("fold %d rows, %d sweeps: %.3f ms" % (ptr.size - 1, sweeps, 1e3 * t))
    _close(W, oracle_lib.fold_rows(ptr, idx, val, H, np.float32(lam), sweeps), "zero start, %d sweeps" % sweeps)
    assert np.all(W[1] == 0.0)
    W0 = (np.random.default_rng(2).normal(0, 0.3, W.shape)).astype(np.float32)
    Ww, _ = fold.rows(ptr, idx, val, sweeps, W0=W0)
    _close(Ww, oracle_lib.fold_rows(ptr, idx, val, H, np.float32(lam), sweeps, W0=W0), "warm start, %d sweeps" % sweeps)
    fold.close()


def test_fold_rows_very_long_row(oracle_lib):
    """A row above 16384 ratings: registers for the first 16384, the rest
    streamed through the global residual buffer."""
    F = _gpu()
    rng = np.random.default_rng(6)
    n, k = 20000, 8
    deg = [17000, 5, 0, 1100]
    ptr = np.concatenate([[0], np.cumsum(deg)]).astype(np.uint32)
    idx = np.concatenate([np.sort(rng.choice(n, size=d, replace=False)) for d in deg]).astype(np.uint32)
    val = rng.integers(1, 6, size=idx.size).astype(np.float32)
    H = (rng.random((k, n)) * 0.5).astype(np.float32)
    fold = F.FoldIn(H, 0.05)
    W, _ = fold.rows(ptr, idx, val, 3)
    _close(W, oracle_lib.fold_rows(ptr, idx, val, H, np.float32(0.05), 3), "very long row")
    np.testing.assert_allclose(fold.features(ptr, idx, val), oracle_lib.fold_features(ptr, idx, val, H), rtol=1e-6,
                               atol=1e-7)
    fold.close()


def test_fold_rows_reach_ridge_solution():
    F = _gpu()
    ptr, idx, val, H = _rows(3, 60, 500, 8)
    lam = 0.1
    fold = F.FoldIn(H, lam)
    W, _ = fold.rows(ptr, idx, val, 300)
    for i in range(ptr.size - 1):
        s, e = int(ptr[i]), int(ptr[i + 1])
        if e == s:
            continue
        Ho = H[:, idx[s:e]].astype(np.float64)
        ref = np.linalg.solve(Ho @ Ho.T + lam * (e - s) * np.eye(H.shape[0]), Ho @ val[s:e].astype(np.float64))
        np.testing.assert_allclose(W[i], ref, rtol=1e-3, atol=1e-4)
    fold.close()


def _split(P, dims):
    nin, hid, out = dims
    sizes = [hid * nin, hid, hid * hid, hid, out * hid, out]
    return np.split(np.asarray(P), np.cumsum(sizes)[:-1])


@pytest.mark.parametrize("rows", [96, 1024])  # 1024: the weight gradients split K over slices
def test_mlp_forward_and_gradients_match_oracle(oracle_lib, rows):
    F = _gpu()
    dims = (82, 256, 40)
    mlp = F.MLP(*dims, seed=7)
    P = mlp.get_params().astype(np.float64)
    rng = np.random.default_rng(4)
    Xn = rng.normal(0, 1, (rows, dims[0])).astype(np.float32)
    Tn = rng.normal(0, 1, (rows, dims[2])).astype(np.float32)
    loss, G = mlp.loss_grad(Xn, Tn)
    rloss, RG = oracle_lib.mlp_backward(P, Xn.astype(np.float64), Tn.astype(np.float64), dims)
    assert abs(loss - rloss) <= 1e-5 * abs(rloss)
    for q, (g, r) in enumerate(zip(_split(G, dims), _split(RG, dims))):
        f = np.linalg.norm(g - r) / np.linalg.norm(r)
        _say("gradient tensor %d: relative %.2e" % (q, f))
        assert f <= 1e-4, "gradient tensor %d: %.3g" % (q, f)
    # forward with standardisation, FP32 and bfloat16 weights
    X = rng.normal(3, 2, (300, dims[0])).astype(np.float32)
    mu, sd = X.mean(0), X.std(0)
    omu, osd = rng.normal(0, 1, dims[2]).astype(np.float32), rng.uniform(0.5, 2, dims[2]).astype(np.float32)
    mlp.set_norm(mu, sd, omu, osd)
    Xn64 = (X.astype(np.float64) - mu) / sd.astype(np.float64)
    ref = oracle_lib.mlp_forward(P, Xn64, dims) * osd + omu
    _close(mlp.forward(X), ref, "forward fp32", frob=1e-5, rtol=1e-4)
    Pb = np.array([oracle_lib.bf16(float(x)) for x in P.astype(np.float32)])
    refb = oracle_lib.mlp_forward(Pb, Xn64, dims) * osd + omu
    _close(mlp.forward(X, bf16=True), refb, "forward bf16 weights", frob=1e-4, rtol=1e-3)
    mlp.close()


def test_mlp_training_lowers_loss():
    F = _gpu()
    rng = np.random.default_rng(5)
    X = rng.normal(0, 1, (8192, 12)).astype(np.float32)
    A = rng.normal(0, 0.5, (12, 4))
    T = (np.tanh(X @ A) + 0.1 * np.sin(3 * X[:, :4])).astype(np.float32)
    mlp = F.MLP(12, 64, 4, seed=1)
    mlp.set_norm(X.mean(0), X.std(0), T.mean(0), T.std(0))
    first = mlp.train(X, T, epochs=1, batch=256, lr=1e-3, seed=1)
    last = mlp.train(X, T, epochs=15, batch=256, lr=1e-3, seed=2)
    _say("training loss %.4f -> %.4f" % (first, last))
    assert last < 0.1 * first
    mlp.close()
