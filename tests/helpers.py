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
