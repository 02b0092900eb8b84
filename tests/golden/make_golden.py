"""Write the golden fixtures of tests/golden/ from the compiled reference.

Runs only where the reference checkout is present (it builds and runs
oracle/_ref/ref_harness, i.e. the reference's own cdmf_ref.cpp + tools.cpp).
The fixtures are data: inputs and the reference's outputs.

Cases (k, lambda, outer, inner):
  simple         data/simple of the reference, 10, 0.05, 5, 3 (its CTest flags),
                 initial_col init, H first
  tricky_simple  data/tricky_simple (column 2 empty in train), same flags
  small_ref      the 2,000 x 1,000 / 50,000-entry synthetic case, 10, 0.05, 5, 5,
                 initial_col init, H first
  small_wfirst   the same matrix with W = 0, H ~ U[0, 1/sqrt(10)) and the W-first
                 order, produced by running the reference on R^T with the roles of
                 W and H exchanged (the shallow transpose of src/pmf_util.h:38-53)

usage: python tests/golden/make_golden.py
"""
from __future__ import annotations

import hashlib
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from cdmf_amd import data as D  # noqa: E402
from cdmf_amd import datadir  # noqa: E402

REF = "/root/reference"
HARNESS = os.path.join(ROOT, "oracle", "_ref", "ref_harness")


def sha(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def run_harness(ddir, k, lam, outer, inner, W0=None, H0=None):
    out = tempfile.mkdtemp(prefix="golden_out_")
    cmd = [HARNESS, ddir, out, str(k), repr(lam), str(outer), str(inner)]
    if W0 is not None:
        wf, hf = os.path.join(out, "W0.bin"), os.path.join(out, "H0.bin")
        np.ascontiguousarray(W0, dtype=np.float32).tofile(wf)
        np.ascontiguousarray(H0, dtype=np.float32).tofile(hf)
        cmd += [wf, hf]
    subprocess.run(cmd, check=True, stdout=subprocess.DEVNULL)
    ds = datadir.read_data_dir(ddir)
    rmse = np.array([float(x) for x in open(os.path.join(out, "rmse.txt")).read().split()])
    Ws = [np.fromfile(os.path.join(out, "W_%d.bin" % it), dtype=np.float32).reshape(k, ds.m) for it in range(1, outer + 1)]
    Hs = [np.fromfile(os.path.join(out, "H_%d.bin" % it), dtype=np.float32).reshape(k, ds.n) for it in range(1, outer + 1)]
    res_csr = np.fromfile(os.path.join(out, "res_csr.bin"), dtype=np.float32)
    res_csc = np.fromfile(os.path.join(out, "res_csc.bin"), dtype=np.float32)
    return ds, rmse, Ws, Hs, res_csr, res_csc


def dataset_arrays(ds):
    return dict(m=ds.m, n=ds.n, csr_ptr=ds.csr_ptr, csr_idx=ds.csr_idx, csr_val=ds.csr_val, csc_ptr=ds.csc_ptr,
                csc_idx=ds.csc_idx, csc_val=ds.csc_val, test_row=ds.test_row, test_col=ds.test_col,
                test_val=ds.test_val)


def main():
    if not os.path.isdir(REF):
        sys.exit("reference checkout not found; fixtures can only be regenerated where it is present")
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], check=True)
    work = tempfile.mkdtemp(prefix="golden_data_")

    for name in ("simple", "tricky_simple"):
        ddir = os.path.join(work, name)
        datadir.convert_text_dir(os.path.join(REF, "data", name), ddir)
        k, lam, outer, inner = 10, 0.05, 5, 3
        ds, rmse, Ws, Hs, rcsr, rcsc = run_harness(ddir, k, lam, outer, inner)
        np.savez_compressed(os.path.join(HERE, name + ".npz"), k=k, lam=lam, outer=outer, inner=inner, rmse=rmse,
                            W=np.stack(Ws), H=np.stack(Hs), res_csr=rcsr, res_csc=rcsc, **dataset_arrays(ds))
        print(name, "rmse", rmse)

    ds = D.small()
    k, lam, outer, inner = D.SMALL_SOLVE.k, D.SMALL_SOLVE.lam, D.SMALL_SOLVE.outer, D.SMALL_SOLVE.inner
    ddir = os.path.join(work, "small")
    datadir.write_data_dir(ddir, ds)
    _, rmse, Ws, Hs, rcsr, rcsc = run_harness(ddir, k, lam, outer, inner)
    np.savez_compressed(os.path.join(HERE, "small_ref.npz"), k=k, lam=lam, outer=outer, inner=inner, rmse=rmse,
                        W=Ws[-1], H=Hs[-1], W_sha=[sha(w) for w in Ws], H_sha=[sha(h) for h in Hs],
                        res_csr_sha=sha(rcsr), res_csc_sha=sha(rcsc), data_sha=dataset_sha(ds))
    print("small_ref rmse", rmse)

    W0, H0 = D.small_init(ds, k)
    tdir = os.path.join(work, "small_T")
    datadir.write_data_dir(tdir, ds.transpose())
    # reference on R^T: its "W" is our H (random), its "H" is our W (zeros)
    _, rmse, Hs_t, Ws_t, rcsc_t, rcsr_t = run_harness(tdir, k, lam, outer, inner, W0=H0, H0=W0)
    np.savez_compressed(os.path.join(HERE, "small_wfirst.npz"), k=k, lam=lam, outer=outer, inner=inner, rmse=rmse,
                        W=Ws_t[-1], H=Hs_t[-1], W_sha=[sha(w) for w in Ws_t], H_sha=[sha(h) for h in Hs_t],
                        res_csr_sha=sha(rcsr_t), res_csc_sha=sha(rcsc_t), data_sha=dataset_sha(ds),
                        H0_sha=sha(H0))
    print("small_wfirst rmse", rmse)


def dataset_sha(ds) -> str:
    h = hashlib.sha256()
    for a in (ds.csr_ptr, ds.csr_idx, ds.csr_val, ds.csc_ptr, ds.csc_idx, ds.csc_val, ds.test_row, ds.test_col,
              ds.test_val):
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


if __name__ == "__main__":
    main()
