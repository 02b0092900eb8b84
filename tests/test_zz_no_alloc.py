"""ccd_fold_rows_noalloc (include/cdmf_foldin.h): the fold-in of a batch of
rows with no device allocation, its batch staged in the library's static
scratch.  Checked through the C ABI against ccd_fold_rows, the allocating
entry point with the same kernels: bit for bit, on two calls in a row (the
second with fewer rows and ratings, so the scratch holds the first call's
leftovers past the second call's data), with and without start factors,
with rows longer than the short-row bound (a block each, residuals in the
scratch's buffer), and a batch over the limit is refused with a message.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _batch(rng, rows, n, mean_deg, long_rows=()):
    deg = np.maximum(1, rng.poisson(mean_deg, size=rows))
    for r, d in long_rows:
        deg[r] = d
    ptr = np.zeros(rows + 1, np.uint32)
    ptr[1:] = np.cumsum(deg)
    idx = np.concatenate([np.sort(rng.choice(n, size=int(d), replace=False)) for d in deg]).astype(np.uint32)
    val = rng.integers(1, 6, size=idx.size).astype(np.float32)
    return ptr, idx, val


def test_fold_rows_noalloc_equals_fold_rows():
    from cdmf_amd import foldin as F
    from cdmf_amd import solver as S
    if S.device_count() < 1:
        pytest.fail("no HIP device visible: GPU tests need the MI355X")
    rng = np.random.default_rng(17)
    k, n, lam = 40, 5000, 0.05
    H = (rng.random((k, n), dtype=np.float32) * 0.3).astype(np.float32)
    big = _batch(rng, 4096, n, 60.0, long_rows=((5, 3000), (4000, 1500)))
    small = _batch(rng, 700, n, 25.0, long_rows=((3, 2048),))
    W0 = (rng.random((700, k), dtype=np.float32) * 0.1).astype(np.float32)
    fo = F.FoldIn(H, lam)
    try:
        for label, (ptr, idx, val), w0, sweeps in (("4096 rows", big, None, 3), ("700 rows, W0", small, W0, 2),
                                                   ("700 rows", small, None, 4)):
            print("case %s: %d ratings, %d sweeps" % (label, idx.size, sweeps), flush=True)
            ref, _ = fo.rows(ptr, idx, val, sweeps, W0=w0)
            got, _ = fo.rows_noalloc(ptr, idx, val, sweeps, W0=w0)
            assert np.all(np.isfinite(ref))
            assert np.array_equal(got, ref), "%s: max |diff| %.3g" % (label, float(np.max(np.abs(got - ref))))
        ptr, idx, val = _batch(rng, 4097, n, 2.0)
        with pytest.raises(S.CCDError) as e:
            fo.rows_noalloc(ptr, idx, val, 1)
        assert "at most 4096 rows" in str(e.value)
    finally:
        fo.close()
