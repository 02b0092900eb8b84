"""The reference's on-disk data-directory format, read and written with numpy.

A data directory holds
  meta_modified_all   text: "m n" / nnz / 9 train file names / nnz_test / 3 test file names
                      (read by the reference loader, src/tools.cpp:271-347)
  meta                text: "m n" / "nnz train.txt" / "nnz_test test.txt"
                      (read by the -p text-test RMSE, src/tools.cpp:489-541)
  R_train_coo.{data,row,col}.bin    COO triplets of the training set
  R_train_csr.{indptr,indices,data}.bin, R_train_csc.{indptr,indices,data}.bin
  R_test_coo.{data,row,col}.bin     COO test set
  train.txt / test.txt              "i j v" lines, 1-based
Binary files are int32 indices / float32 values, 0-based
(data/simple/simple_compress.py:79-246 writes this layout with scipy).

Differences from the reference converter, all deliberate:
  * the compressed pointers always have m+1 / n+1 entries (scipy infers the
    shape from the largest index, so the reference files can be short;
    `read_data_dir` pads a short pointer with nnz, as the reference loader
    effectively does when its read runs off the end of the file);
  * duplicate (i, j) training pairs are summed, as scipy's tocsr() does, and
    the stored nnz is the number of distinct pairs;
  * no numpy.float (removed in numpy 1.24), so it runs on current numpy.
"""
from __future__ import annotations

import os
from dataclasses import dataclass

import numpy as np

TRAIN_FILES = (
    "R_train_coo.data.bin",
    "R_train_coo.row.bin",
    "R_train_coo.col.bin",
    "R_train_csr.indptr.bin",
    "R_train_csr.indices.bin",
    "R_train_csr.data.bin",
    "R_train_csc.indptr.bin",
    "R_train_csc.indices.bin",
    "R_train_csc.data.bin",
)
TEST_FILES = ("R_test_coo.data.bin", "R_test_coo.row.bin", "R_test_coo.col.bin")


@dataclass
class Dataset:
    """Training matrix in both compressed layouts plus the COO test set.

    Mirrors SparseMatrix (src/pmf_util.h:13-114) and TestData
    (src/pmf_util.h:117-161): uint32 pointers/indices, float32 values.
    """

    m: int
    n: int
    csr_ptr: np.ndarray
    csr_idx: np.ndarray
    csr_val: np.ndarray
    csc_ptr: np.ndarray
    csc_idx: np.ndarray
    csc_val: np.ndarray
    test_row: np.ndarray
    test_col: np.ndarray
    test_val: np.ndarray

    @property
    def nnz(self) -> int:
        return int(self.csr_idx.size)

    @property
    def nnz_test(self) -> int:
        return int(self.test_row.size)

    def copy(self) -> "Dataset":
        return Dataset(self.m, self.n, *(a.copy() for a in (
            self.csr_ptr, self.csr_idx, self.csr_val, self.csc_ptr, self.csc_idx,
            self.csc_val, self.test_row, self.test_col, self.test_val)))

    def transpose(self) -> "Dataset":
        """R^T: the shallow transpose of src/pmf_util.h:38-53 (CSR and CSC swap)."""
        return Dataset(self.n, self.m, self.csc_ptr, self.csc_idx, self.csc_val,
                       self.csr_ptr, self.csr_idx, self.csr_val,
                       self.test_col, self.test_row, self.test_val)


def compress(m: int, n: int, rows: np.ndarray, cols: np.ndarray, vals: np.ndarray):
    """COO -> (CSR, CSC) with duplicates summed and indices sorted, like
    scipy's coo_matrix.tocsr()/tocsc().  Sums are taken in float64 and then
    rounded to float32, as the reference converter does."""
    rows = np.asarray(rows, dtype=np.int64)
    cols = np.asarray(cols, dtype=np.int64)
    vals = np.asarray(vals, dtype=np.float64)
    if rows.size and (rows.min() < 0 or rows.max() >= m or cols.min() < 0 or cols.max() >= n):
        raise ValueError("index out of range for shape (%d, %d)" % (m, n))
    key = rows * n + cols
    order = np.argsort(key, kind="stable")
    key_s = key[order]
    vals_s = vals[order]
    if key_s.size:
        first = np.ones(key_s.size, dtype=bool)
        first[1:] = key_s[1:] != key_s[:-1]
        starts = np.flatnonzero(first)
        ukey = key_s[starts]
        uval = np.add.reduceat(vals_s, starts)
    else:
        ukey = key_s
        uval = vals_s
    r = (ukey // n).astype(np.int64)
    c = (ukey % n).astype(np.int64)
    csr_ptr = np.zeros(m + 1, dtype=np.int64)
    np.add.at(csr_ptr, r + 1, 1)
    csr_ptr = np.cumsum(csr_ptr)
    csr_idx = c.astype(np.uint32)
    csr_val = uval.astype(np.float32)
    # CSC: stable sort by column keeps rows ascending inside each column
    corder = np.argsort(c, kind="stable")
    csc_ptr = np.zeros(n + 1, dtype=np.int64)
    np.add.at(csc_ptr, c + 1, 1)
    csc_ptr = np.cumsum(csc_ptr)
    csc_idx = r[corder].astype(np.uint32)
    csc_val = uval[corder].astype(np.float32)
    return (csr_ptr.astype(np.uint32), csr_idx, csr_val, csc_ptr.astype(np.uint32), csc_idx, csc_val)


def from_triplets(m: int, n: int, train, test) -> Dataset:
    """Build a Dataset from 0-based (rows, cols, vals) train and test triplets."""
    csr = compress(m, n, *train)
    trow, tcol, tval = test
    return Dataset(m, n, *csr, np.asarray(trow, dtype=np.uint32), np.asarray(tcol, dtype=np.uint32),
                   np.asarray(tval, dtype=np.float32))


def read_triplet_text(path: str):
    """Read an 'i<TAB>j<TAB>r' (or space separated) 1-based triplet file."""
    rows, cols, vals = [], [], []
    with open(path) as f:
        for line in f:
            parts = line.split()
            if len(parts) < 3:
                continue
            rows.append(int(parts[0]) - 1)
            cols.append(int(parts[1]) - 1)
            vals.append(float(parts[2]))
    return (np.array(rows, dtype=np.int64), np.array(cols, dtype=np.int64), np.array(vals, dtype=np.float64))


def _fmt_rating(v: float) -> str:
    # the reference writes str(numpy.float64), e.g. "4.0"
    return repr(float(v))


def write_data_dir(path: str, ds: Dataset, train_coo=None) -> None:
    """Write `ds` in the reference's data-directory layout.

    train_coo: optional (rows, cols, vals) triplets in file order for the
    R_train_coo.* files and train.txt; defaults to the CSR order.
    """
    os.makedirs(path, exist_ok=True)
    if train_coo is None:
        r = np.repeat(np.arange(ds.m, dtype=np.int64), np.diff(ds.csr_ptr.astype(np.int64)))
        train_coo = (r, ds.csr_idx.astype(np.int64), ds.csr_val.astype(np.float64))
    tr, tc, tv = (np.asarray(a) for a in train_coo)
    j = lambda name: os.path.join(path, name)  # noqa: E731
    np.asarray(tv, dtype=np.float32).tofile(j("R_train_coo.data.bin"))
    np.asarray(tr, dtype=np.int32).tofile(j("R_train_coo.row.bin"))
    np.asarray(tc, dtype=np.int32).tofile(j("R_train_coo.col.bin"))
    ds.csr_ptr.astype(np.int32).tofile(j("R_train_csr.indptr.bin"))
    ds.csr_idx.astype(np.int32).tofile(j("R_train_csr.indices.bin"))
    ds.csr_val.astype(np.float32).tofile(j("R_train_csr.data.bin"))
    ds.csc_ptr.astype(np.int32).tofile(j("R_train_csc.indptr.bin"))
    ds.csc_idx.astype(np.int32).tofile(j("R_train_csc.indices.bin"))
    ds.csc_val.astype(np.float32).tofile(j("R_train_csc.data.bin"))
    ds.test_val.astype(np.float32).tofile(j("R_test_coo.data.bin"))
    ds.test_row.astype(np.int32).tofile(j("R_test_coo.row.bin"))
    ds.test_col.astype(np.int32).tofile(j("R_test_coo.col.bin"))
    with open(j("meta_modified_all"), "w") as f:
        f.write("%d %d\n%d\n" % (ds.m, ds.n, ds.nnz))
        f.write("\n".join(TRAIN_FILES) + "\n")
        f.write("%d\n" % ds.nnz_test)
        f.write("\n".join(TEST_FILES) + "\n")
    with open(j("meta"), "w") as f:
        f.write("%d %d\n%d train.txt\n%d test.txt\n" % (ds.m, ds.n, len(tr), ds.nnz_test))
    with open(j("train.txt"), "w") as f:
        for a, b, c in zip(tr, tc, tv):
            f.write("%d %d %s\n" % (a + 1, b + 1, _fmt_rating(c)))
    with open(j("test.txt"), "w") as f:
        for a, b, c in zip(ds.test_row, ds.test_col, ds.test_val):
            f.write("%d %d %s\n" % (a + 1, b + 1, _fmt_rating(c)))


def _read_ptr(fname: str, count: int, nnz: int) -> np.ndarray:
    raw = np.fromfile(fname, dtype=np.int32)
    out = np.full(count, nnz, dtype=np.uint32)
    k = min(count, raw.size)
    out[:k] = raw[:k].astype(np.uint32)
    return out


class DataDirError(RuntimeError):
    pass


def read_data_dir(path: str) -> Dataset:
    """Load a data directory (src/tools.cpp:271-347, src/pmf_util.h:80-108)."""
    meta = os.path.join(path, "meta_modified_all")
    if not os.path.exists(meta):
        raise DataDirError("Can't open meta input file.")
    with open(meta) as f:
        toks = f.read().split()
    try:
        m, n, nnz = int(toks[0]), int(toks[1]), int(toks[2])
        names = toks[3:12]
        nnz_test = int(toks[12])
        tnames = toks[13:16]
    except (IndexError, ValueError) as e:
        raise DataDirError("malformed meta_modified_all: %s" % e)
    if len(names) != 9 or len(tnames) != 3:
        raise DataDirError("malformed meta_modified_all")
    j = lambda name: os.path.join(path, name)  # noqa: E731

    def arr(name, dtype, count):
        a = np.fromfile(j(name), dtype=dtype)
        if a.size < count:
            raise DataDirError("FREAD read %d, needed %d (%s)" % (a.size, count, name))
        return a[:count]

    csr_ptr = _read_ptr(j(names[3]), m + 1, nnz)
    csr_idx = arr(names[4], np.int32, nnz).astype(np.uint32)
    csr_val = arr(names[5], np.float32, nnz).copy()
    csc_ptr = _read_ptr(j(names[6]), n + 1, nnz)
    csc_idx = arr(names[7], np.int32, nnz).astype(np.uint32)
    csc_val = arr(names[8], np.float32, nnz).copy()
    test_val = arr(tnames[0], np.float32, nnz_test).copy()
    test_row = arr(tnames[1], np.int32, nnz_test).astype(np.uint32)
    test_col = arr(tnames[2], np.int32, nnz_test).astype(np.uint32)
    return Dataset(m, n, csr_ptr, csr_idx, csr_val, csc_ptr, csc_idx, csc_val, test_row, test_col, test_val)


def convert_text_dir(src_dir: str, dst_dir: str) -> Dataset:
    """Convert a reference text data set (meta + train/test triplet files,
    e.g. data/simple) into the binary data-directory layout."""
    with open(os.path.join(src_dir, "meta")) as f:
        toks = f.read().split()
    m, n = int(toks[0]), int(toks[1])
    train_name, test_name = toks[3], toks[5]
    tr = read_triplet_text(os.path.join(src_dir, train_name))
    te = read_triplet_text(os.path.join(src_dir, test_name))
    ds = from_triplets(m, n, tr, te)
    write_data_dir(dst_dir, ds, train_coo=tr)
    return ds
