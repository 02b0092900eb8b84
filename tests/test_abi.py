"""CPU-side checks of the C ABI: both libraries load and export every symbol
their headers declare; without a GPU the entry points fail cleanly."""
import ctypes as C
import os
import re

import pytest

from tests.helpers import ROOT


def _declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b((?:ccd|cdmf)_[a-z0-9_]+)\s*\(", text)))


@pytest.mark.parametrize("header,lib", [("cdmf_hip.h", "libcdmf_hip.so"), ("cdmf_data.h", "libcdmf_data.so"),
                                        ("cdmf_foldin.h", "libcdmf_hip.so")])
def test_library_exports_every_declared_symbol(header, lib):
    names = _declared(header)
    assert len(names) >= 8
    L = C.CDLL(os.path.join(ROOT, "cdmf_amd", "_lib", lib))
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_python_binding_covers_the_header():
    from cdmf_amd import foldin, solver
    assert sorted(solver.ABI_SYMBOLS) == _declared("cdmf_hip.h")
    assert sorted(foldin.FOLDIN_SYMBOLS) == _declared("cdmf_foldin.h")
    foldin._lib()  # every declared signature binds


def test_no_device_fails_cleanly():
    from cdmf_amd import solver
    if solver.device_count() > 0:
        pytest.skip("a GPU is visible")
    with pytest.raises(solver.CCDError):
        solver.Solver(0)
    assert solver.hip_lib().ccd_abi_version() == 6


def test_params_default_are_the_reference_defaults():
    from cdmf_amd import solver
    p = solver.CParams()
    solver.hip_lib().ccd_params_default(C.byref(p))
    # src/pmf.h:4-26
    assert (p.k, round(p.lambda_, 6), p.maxiter, p.maxinneriter, p.variant) == (10, 0.1, 5, 5, 1)
    assert p.zero_first == 1 and p.order == 0
