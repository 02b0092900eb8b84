"""CPU check of the tiled layout's host tables (cdmf_amd/csrc/tile_layout.cpp),
built with the address and undefined-behaviour sanitizers
(tests/native/tile_check.cpp): every tiled layout is decoded back to its
compressed layout, and a restatement of the tiled sweep's per-lane, per-tile
and cross-tile bookkeeping must reproduce a direct pass over the layout
exactly, for sigma 8, 16 and 32, plain and split, with 16-bit, 3-byte and
32-bit index planes.

Cases: the 2,000 x 1,000 matrix (both layouts), the same with every 3rd row
and 5th column emptied (segment tables), a tall matrix with 70,000 rows
(3-byte CSC indices, split at 40,704 like the library on MI355X), the
90,000 x 2,500 power-law matrix of the split parity test (split CSC copy),
the reference's stress shape b (a dense row and column of 1,024,000
ratings: both layouts split, 3-byte indices, spans of about 1,000 tiles),
and edge shapes (one row, one column, a 3 x 2 matrix, a full 40 x 40 block).
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

from tests.helpers import DEGENERATE, ROOT, degenerate

SRC = os.path.join(ROOT, "tests", "native", "tile_check.cpp")
LAYOUT = os.path.join(ROOT, "cdmf_amd", "csrc", "tile_layout.cpp")
INC = os.path.join(ROOT, "cdmf_amd", "csrc")
SPLIT_AT = 40704  # the LDS prefix of a plain sweep on MI355X (160 KiB per CU, one block): the library's split


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    out = str(tmp_path_factory.mktemp("tile_check") / "tile_check")
    subprocess.run(["g++", "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
                    "-fno-sanitize-recover=undefined", "-I" + INC, SRC, LAYOUT, "-o", out], check=True)
    return out


def _layouts(name):
    from cdmf_amd import data as D
    from cdmf_amd import datadir
    if name == "small":
        ds = D.small()
    elif name == "holey":
        s = D.small()
        rows = np.repeat(np.arange(s.m), np.diff(s.csr_ptr.astype(np.int64)))
        keep = (rows % 3 != 0) & (s.csr_idx % 5 != 0)
        ds = datadir.from_triplets(s.m, s.n, (rows[keep], s.csr_idx[keep], s.csr_val[keep]),
                                   (s.test_row, s.test_col, s.test_val))
    elif name == "tall":
        rng = np.random.default_rng(11)
        m, n = 70000, 400
        rows = np.sort(rng.choice(m, size=30000, replace=False)).astype(np.int64)
        r = np.repeat(rows, 12)
        c = rng.integers(0, n, size=r.size)
        key = np.unique(r * n + c)
        r, c = key // n, key % n
        v = np.ones(r.size, np.float32)
        ds = datadir.from_triplets(m, n, (r, c, v), (r[:10], c[:10], v[:10]))
    elif name in DEGENERATE:
        ds = degenerate(name)
    elif name == "powerlaw":
        ds = D.gen_powerlaw(m=90000, n=2500, nnz_train=450000, nnz_test=20000, max_row_degree=2500, rating_kind=1,
                            seed=21)
    else:
        ds = D.gen_stress("b", seed=0)
    # (label, rows, gathered dimension, ptr, idx)
    return [("csr", ds.m, ds.n, ds.csr_ptr, ds.csr_idx), ("csc", ds.n, ds.m, ds.csc_ptr, ds.csc_idx)]


def _write(path, rows, gdim, ptr, idx):
    with open(path, "wb") as f:
        np.array([rows, gdim], np.int64).tofile(f)
        np.array([idx.size], np.uint64).tofile(f)
        np.ascontiguousarray(ptr, np.uint32).tofile(f)
        np.ascontiguousarray(idx, np.uint32).tofile(f)


@pytest.mark.parametrize("sigma", [8, 16, 32])
@pytest.mark.parametrize("name", ["small", "holey", "tall", "powerlaw", "stress_b"] + DEGENERATE)
def test_tile_tables_decode_to_the_layout(checker, tmp_path, name, sigma):
    for label, rows, gdim, ptr, idx in _layouts(name):
        path = str(tmp_path / ("%s_%s.bin" % (name, label)))
        _write(path, rows, gdim, ptr, idx)
        splits = [0] + ([SPLIT_AT] if gdim > SPLIT_AT else [])
        for split_at in splits:
            env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=0",
                       UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
            r = subprocess.run([checker, path, str(sigma), str(split_at)], capture_output=True, text=True,
                               timeout=600, env=env)
            print(name, label, r.stdout.strip(), r.stderr[-3000:], flush=True)
            assert r.returncode == 0, "%s %s sigma=%d split_at=%d" % (name, label, sigma, split_at)
            assert r.stdout.startswith("OK"), r.stdout
