"""Shared test helpers (fixture loading)."""
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")


def load_golden(name):
    return dict(np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False))


def golden_dataset(g):
    from cdmf_amd.datadir import Dataset
    return Dataset(int(g["m"]), int(g["n"]), g["csr_ptr"], g["csr_idx"], g["csr_val"], g["csc_ptr"], g["csc_idx"],
                   g["csc_val"], g["test_row"], g["test_col"], g["test_val"])


DEGENERATE = ["one_row", "one_col", "tiny", "dense"]


def degenerate(name):
    """Edge shapes: one row of 2,500 ratings over 3,000 columns (a row over
    three tiles; 500 empty columns), its transpose, a 3 x 2 matrix of four
    ratings (one padded tile) and a full 40 x 40 block.  Ratings are seeded
    integers 1..5; the test entries are positions outside the training set
    (for the full block, every 37th of its own positions)."""
    from cdmf_amd import datadir
    rng = np.random.default_rng({"one_row": 1, "one_col": 2, "tiny": 3, "dense": 4}[name])
    if name in ("one_row", "one_col"):
        pos = rng.permutation(3000)
        tr, te = pos[:2500], pos[2500:2540]
        z = lambda a: np.zeros(a.size, np.int64)
        if name == "one_row":
            m, n, train, test = 1, 3000, (z(tr), tr), (z(te), te)
        else:
            m, n, train, test = 3000, 1, (tr, z(tr)), (te, z(te))
    elif name == "tiny":
        m, n = 3, 2
        train = (np.array([0, 0, 2, 1]), np.array([0, 1, 1, 0]))
        test = (np.array([1, 2]), np.array([1, 0]))
    else:
        m, n = 40, 40
        r, c = np.divmod(np.arange(1600), 40)
        train, test = (r, c), (r[::37], c[::37])
    val = lambda k: rng.integers(1, 6, size=k).astype(np.float32)
    return datadir.from_triplets(m, n, (*train, val(train[0].size)), (*test, val(test[0].size)))
