"""The command-line driver (flags and output lines of the reference's main.cpp).

CPU-side: argument errors and missing data exit the way the reference does
(src/tools.cpp:432-435, 451-454, 276-279; src/main.cpp:60-63).
GPU-side: an end-to-end run on the reference's `simple` data set prints the
per-iteration line, the training-time line the reference's scripts grep
(scripts/gpgpu.sh:24-26) and, with -r 1, golden_compare's verdict against the
reference's own factors.
"""
import os
import re
import subprocess

import numpy as np
import pytest

from tests.helpers import ROOT, golden_dataset, load_golden

CLI = os.path.join(ROOT, "cdmf_amd", "bin", "cdmf")


def run(args, **kw):
    return subprocess.run([CLI, *args], capture_output=True, text=True, timeout=300, **kw)


def test_help_on_missing_data_dir():
    r = run(["-k", "4"])
    assert r.returncode != 0 and "Usage: cdmf" in r.stdout


def test_unknown_option():
    r = run(["-z", "1", "/tmp"])
    assert r.returncode != 0 and "unknown option: -z" in r.stderr


def test_unknown_device_type():
    r = run(["-d", "7", "/tmp"])
    assert r.returncode != 0 and "unknown device type!" in r.stderr


def test_missing_meta(tmp_path):
    r = run([str(tmp_path)])
    assert r.returncode != 0 and "Can't open meta input file." in r.stdout


def test_truncated_binary_aborts(tmp_path):
    from cdmf_amd import datadir
    ds = golden_dataset(load_golden("simple"))
    datadir.write_data_dir(str(tmp_path), ds)
    with open(tmp_path / "R_train_csr.data.bin", "r+b") as f:
        f.truncate(8)
    r = run([str(tmp_path)])
    assert r.returncode != 0 and "FREAD" in r.stderr


def _write_simple(tmp_path):
    from cdmf_amd import datadir
    g = load_golden("simple")
    ds = golden_dataset(g)
    datadir.write_data_dir(str(tmp_path), ds)
    return g, ds


@pytest.mark.gpu
@pytest.mark.parametrize("version", [1, 2, 3])
def test_end_to_end_simple(tmp_path, version):
    """-V 3 is the CSR5 driver's signature (no test set): no per-iteration
    RMSE line, the text-test RMSE of -p 1 only."""
    g, ds = _write_simple(tmp_path)
    k, T = int(g["k"]), int(g["outer"])
    # reference factors of the reference's own cdmf_ref for -r 1
    g["W"][-1].astype(np.float32).tofile(tmp_path / "W_ref.bin")
    g["H"][-1].astype(np.float32).tofile(tmp_path / "H_ref.bin")
    r = run(["-k", str(k), "-t", str(T), "-T", str(int(g["inner"])), "-l", "0.05", "-p", "1", "-r", "1",
             "-V", str(version), str(tmp_path)])
    print(r.stdout[-3000:], r.stderr[-2000:], flush=True)
    assert r.returncode == 0
    assert "[info] Picked Version %d" % version in r.stdout
    lines = re.findall(r"\[-INFO-\] iteration num (\d+) .*RMSE=([0-9.]+)", r.stdout)
    if version == 3:
        assert lines == [] and "[INFO] spmv time" in r.stdout
    else:
        assert [int(a) for a, _ in lines] == list(range(1, T + 1))
        np.testing.assert_allclose([float(b) for _, b in lines], g["rmse"], atol=2e-6)
    assert re.search(r"\[INFO\] OCL Training time: [0-9.]+ s", r.stdout)
    m = re.search(r"\[INFO\] Test RMSE = ([0-9.]+)", r.stdout)
    assert m and abs(float(m.group(1)) - g["rmse"][-1]) < 2e-6
    assert r.stdout.count("Check... PASS!") == 2
