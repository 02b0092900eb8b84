"""CPU checks of the data layer: seeded generators (cdmf_data.h), the
data-directory reader/writer (src/tools.cpp:271-347 format, written by
data/simple/simple_compress.py:79-246 in the reference) and the converter's
COO semantics (duplicates summed, indices sorted, like scipy)."""
import numpy as np
import pytest
import scipy.sparse as sp

from cdmf_amd import data as D
from cdmf_amd import datadir


def _check_compressed(ds):
    for ptr, idx, dim, rows in ((ds.csr_ptr, ds.csr_idx, ds.n, ds.m), (ds.csc_ptr, ds.csc_idx, ds.m, ds.n)):
        assert ptr[0] == 0 and ptr[-1] == ds.nnz and np.all(np.diff(ptr.astype(np.int64)) >= 0)
        assert idx.size == ds.nnz and (idx.size == 0 or idx.max() < dim)
        for i in range(0, rows, max(1, rows // 50)):  # indices strictly ascending inside rows
            seg = idx[ptr[i]:ptr[i + 1]].astype(np.int64)
            assert np.all(np.diff(seg) > 0)
    A = sp.csr_matrix((ds.csr_val, ds.csr_idx, ds.csr_ptr), shape=(ds.m, ds.n))
    B = sp.csc_matrix((ds.csc_val, ds.csc_idx, ds.csc_ptr), shape=(ds.m, ds.n))
    assert (A != B).nnz == 0


@pytest.mark.parametrize("kind", [0, 1, 2, 3])
def test_powerlaw_generator(kind):
    ds = D.gen_powerlaw(m=3000, n=700, nnz_train=90000, nnz_test=9000, max_row_degree=700, rating_kind=kind, seed=9)
    assert ds.nnz == 90000 and ds.nnz_test == 9000
    _check_compressed(ds)
    deg = np.diff(ds.csr_ptr.astype(np.int64))
    assert deg.max() <= 700
    again = D.gen_powerlaw(m=3000, n=700, nnz_train=90000, nnz_test=9000, max_row_degree=700, rating_kind=kind, seed=9,
                           threads=1)
    for a, b in ((ds.csr_idx, again.csr_idx), (ds.csr_val, again.csr_val), (ds.test_val, again.test_val)):
        assert np.array_equal(a, b)  # independent of the thread count
    vals = np.concatenate([ds.csr_val, ds.test_val])
    if kind in (1, 3):  # the 1..5 frequencies .05 .10 .29 .34 .22
        freq = np.array([np.mean(vals == r) for r in range(1, 6)])
        np.testing.assert_allclose(freq, [0.05, 0.10, 0.29, 0.34, 0.22], atol=0.015)
    if kind == 2:
        assert np.all(np.isin(vals, np.arange(1, 11) * 0.5))


def test_planted_ratings_carry_low_rank_structure():
    """A rank-8 model explains much of the planted ratings but nothing of the
    structureless ones: compare the item-mean-centred rank-8 SVD energy."""
    def energy(kind):
        ds = D.gen_powerlaw(m=2000, n=400, nnz_train=200000, nnz_test=0, max_row_degree=400, rating_kind=kind, seed=4)
        A = sp.csr_matrix((ds.csr_val, ds.csr_idx, ds.csr_ptr), shape=(ds.m, ds.n)).toarray()
        M = A != 0
        A = np.where(M, A - A.sum(0) / np.maximum(M.sum(0), 1), 0.0)
        s = np.linalg.svd(A, compute_uv=False)
        return (s[:8] ** 2).sum() / (s ** 2).sum()
    assert energy(3) > energy(1) + 0.1


def test_uniform_generator_distinct_positions():
    ds = D.gen_uniform(500, 300, 20000, 2000, seed=3)
    _check_compressed(ds)
    train = set(zip(np.repeat(np.arange(ds.m), np.diff(ds.csr_ptr.astype(np.int64))).tolist(), ds.csr_idx.tolist()))
    test = set(zip(ds.test_row.tolist(), ds.test_col.tolist()))
    assert len(train) == 20000 and len(test) == 2000 and not (train & test)
    with pytest.raises(RuntimeError):
        D.gen_uniform(10, 10, 50, 0, seed=0, rating_kind=3)  # planted needs the power-law generator


def test_data_dir_round_trip(tmp_path):
    ds = D.gen_powerlaw(m=400, n=120, nnz_train=5000, nnz_test=400, max_row_degree=120, seed=2)
    D.write_dir(str(tmp_path / "a"), ds)
    back = D.load_dir(str(tmp_path / "a"))
    py = datadir.read_data_dir(str(tmp_path / "a"))
    for other in (back, py):
        for name in ("csr_ptr", "csr_idx", "csr_val", "csc_ptr", "csc_idx", "csc_val", "test_row", "test_col",
                     "test_val"):
            assert np.array_equal(getattr(ds, name), getattr(other, name)), name


def test_converter_sums_duplicates_like_scipy():
    rng = np.random.default_rng(0)
    r = rng.integers(0, 30, 400)
    c = rng.integers(0, 20, 400)
    v = rng.integers(1, 6, 400).astype(np.float64)
    csr_ptr, csr_idx, csr_val, csc_ptr, csc_idx, csc_val = datadir.compress(30, 20, r, c, v)
    ref = sp.coo_matrix((v, (r, c)), shape=(30, 20)).tocsr()
    ref.sort_indices()
    assert np.array_equal(csr_ptr, ref.indptr) and np.array_equal(csr_idx, ref.indices)
    np.testing.assert_array_equal(csr_val, ref.data.astype(np.float32))
    refc = sp.coo_matrix((v, (r, c)), shape=(30, 20)).tocsc()
    refc.sort_indices()
    assert np.array_equal(csc_ptr, refc.indptr) and np.array_equal(csc_idx, refc.indices)


def test_batch_specs_are_valid_powerlaw_shapes():
    specs = D.batch_specs()
    assert len(specs) == 64
    for bp in specs:
        assert bp.nnz <= bp.m * bp.n // 2 and bp.nnz >= 2 * bp.m - 2
        assert bp.solve.k in (8, 16, 32, 64) and bp.solve.lam in (0.01, 0.05, 0.1)
    small = D.batch_specs(count=3, scale=0.01)
    for bp in small:  # every drawn shape generates
        ds = D.batch_dataset(bp)
        assert ds.nnz + ds.nnz_test == bp.nnz


@pytest.mark.parametrize("shape", ["a", "b", "c"])
def test_stress_shapes_follow_the_reference_generators(shape):
    """cdmf_gen_stress against the structure of data/{a,b,c}/{a,b,c}.py:
    sizes, the dense row / column, two ratings on every other row (the fixed
    column and one drawn from the stated range), train/100 test entries
    apart from the training ratings, ratings in 1..5, and the same matrix
    for the same seed."""
    ds = D.gen_stress(shape, seed=0)
    _check_compressed(ds)
    m, n = {"a": (64000, 150), "b": (1024000, 1024000), "c": (1024000, 1024000)}[shape]
    assert (ds.m, ds.n) == (m, n)
    ntrain = n + 2 * (m - 1) if shape in "ab" else 2 * m
    assert ds.nnz == ntrain and ds.nnz_test == ntrain // 100
    rdeg = np.diff(ds.csr_ptr.astype(np.int64))
    cdeg = np.diff(ds.csc_ptr.astype(np.int64))
    dense_row = {"a": 0, "b": m - 1, "c": None}[shape]
    dense_col = {"a": 0, "b": n - 1, "c": 0}[shape]
    assert cdeg[dense_col] == m
    sparse = np.ones(m, bool)
    if dense_row is not None:
        assert rdeg[dense_row] == n
        sparse[dense_row] = False
    assert np.all(rdeg[sparse] == 2)
    rows = np.repeat(np.arange(m), rdeg)
    drawn = ds.csr_idx[(rows != (dense_row if dense_row is not None else -1)) & (ds.csr_idx != dense_col)]
    lo, hi = {"a": (1, n), "b": (0, n - 1), "c": (1, n)}[shape]
    assert drawn.size == sparse.sum() and drawn.min() >= lo and drawn.max() < hi
    train = set((rows.astype(np.int64) * n + ds.csr_idx).tolist())
    test = set((ds.test_row.astype(np.int64) * n + ds.test_col).tolist())
    assert len(test) == ds.nnz_test and not (train & test)
    vals = np.concatenate([ds.csr_val, ds.test_val])
    assert np.all(np.isin(vals, [1, 2, 3, 4, 5]))
    again = D.gen_stress(shape, seed=0)
    assert np.array_equal(again.csr_idx, ds.csr_idx) and np.array_equal(again.csr_val, ds.csr_val)
    assert np.array_equal(again.test_col, ds.test_col)
