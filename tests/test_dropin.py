"""Drop-in check: the reference's own main.cpp, tools.cpp and cdmf_ref.cpp,
linked with this project's adapter (cdmf_amd/csrc/host/cdmf_hip.cpp) and
libcdmf_hip.so in place of cdmf_ocl_1.cpp / cdmf_ocl_2.cpp / cdmf_csr5.cpp
(oracle/Makefile target `refmain`, built only where the reference checkout
exists).  The run uses the reference's CTest flags (CMakeLists.txt:61-67)
and its own -r 1 comparison against its OpenMP solver."""
import os
import re
import subprocess

import numpy as np
import pytest

from tests.helpers import ROOT, golden_dataset, load_golden

REFMAIN = os.path.join(ROOT, "oracle", "_ref", "cdmf_refmain")

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not os.path.exists(REFMAIN), reason="oracle/_ref/cdmf_refmain not built")]


@pytest.mark.parametrize("name", ["simple", "tricky_simple"])
@pytest.mark.parametrize("version", [1, 2, 3])
def test_reference_main_with_hip_backend(tmp_path, name, version):
    """-V 1 and -V 2 (cdmf_ocl_01/02, with a test set) print one RMSE line
    per outer iteration; -V 3 (cdmf_csr5, src/cdmf.h:10, main.cpp:53-57)
    takes no test set, so only -p 1's text-test RMSE (tools.cpp:489-541)
    reports its result.  Either way the reference's -r 1 solve follows and
    golden_compare must pass for W and H."""
    from cdmf_amd import datadir
    g = load_golden(name)
    datadir.write_data_dir(str(tmp_path), golden_dataset(g))
    args = [REFMAIN, "-k", "10", "-t", "5", "-T", "3", "-l", "0.05", "-p", "1", "-r", "1", "-q", "0", "-P", "0",
            "-d", "1", "-V", str(version), "-nThreadsPerBlock", "2", str(tmp_path)]
    print("running", " ".join(args), flush=True)
    r = subprocess.run(args, capture_output=True, text=True, timeout=300)
    print(r.stdout[-4000:], r.stderr[-2000:], flush=True)
    assert r.returncode == 0
    # the HIP solver's per-iteration lines (V1, V2), then the reference OpenMP solver's
    rm = [float(x) for x in re.findall(r"RMSE=([0-9.]+) time", r.stdout)]
    hip_lines = 5 if version in (1, 2) else 0
    assert len(rm) == hip_lines + 5
    np.testing.assert_allclose(rm[:hip_lines], g["rmse"][:hip_lines], atol=2e-6)
    np.testing.assert_allclose(rm[hip_lines:], g["rmse"], atol=2e-6)
    # -p 1: the text-test RMSE of the HIP result, then of the OpenMP result
    text = [float(x) for x in re.findall(r"\[INFO\] Test RMSE = ([0-9.]+)", r.stdout)]
    assert len(text) == 2
    np.testing.assert_allclose(text, [g["rmse"][-1]] * 2, atol=2e-6)
    assert "[info] Picked Version %d" % version in r.stdout
    assert r.stdout.count("Check... PASS!") == 2
    assert "OCL Training time" in r.stdout
    if version == 3:
        assert "[INFO] spmv time" in r.stdout
