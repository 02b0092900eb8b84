"""ctypes bindings of the two native libraries (include/cdmf_data.h, include/cdmf_hip.h).

The libraries are loaded from cdmf_amd/_lib/ (or the directory named by the
CDMF_AMD_LIBDIR environment variable, for comparing two builds); nothing here
falls back to a Python or CPU implementation.  A missing library raises
NativeLibraryMissing naming the file and the build command.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBDIR = os.environ.get("CDMF_AMD_LIBDIR") or os.path.join(HERE, "_lib")


class NativeLibraryMissing(RuntimeError):
    pass


def _open(name: str) -> C.CDLL:
    path = os.path.join(LIBDIR, name)
    if not os.path.exists(path):
        raise NativeLibraryMissing(
            "%s is not built; run `python -c 'import __graft_entry__ as g; g.build()'`" % path)
    return C.CDLL(path, mode=C.RTLD_GLOBAL)


u32p = C.POINTER(C.c_uint32)
f32p = C.POINTER(C.c_float)
f64p = C.POINTER(C.c_double)


class CDataset(C.Structure):
    _fields_ = [
        ("m", C.c_int64), ("n", C.c_int64),
        ("nnz", C.c_uint64), ("nnz_test", C.c_uint64),
        ("max_row_nnz", C.c_uint64), ("max_col_nnz", C.c_uint64),
        ("csr_ptr", u32p), ("csr_idx", u32p), ("csr_val", f32p),
        ("csc_ptr", u32p), ("csc_idx", u32p), ("csc_val", f32p),
        ("test_row", u32p), ("test_col", u32p), ("test_val", f32p),
    ]


class CGenSpec(C.Structure):
    _fields_ = [
        ("m", C.c_int64), ("n", C.c_int64),
        ("nnz_train", C.c_uint64), ("nnz_test", C.c_uint64),
        ("max_row_degree", C.c_uint32),
        ("row_alpha", C.c_double), ("col_exponent", C.c_double),
        ("rating_kind", C.c_int32), ("threads", C.c_int32),
        ("seed", C.c_uint64),
    ]


_data = None


def data_lib() -> C.CDLL:
    global _data
    if _data is None:
        lib = _open("libcdmf_data.so")
        lib.cdmf_gen_powerlaw.argtypes = [C.POINTER(CGenSpec), C.POINTER(CDataset)]
        lib.cdmf_gen_uniform.argtypes = [C.c_int64, C.c_int64, C.c_uint64, C.c_uint64, C.c_int32, C.c_uint64,
                                         C.POINTER(CDataset)]
        lib.cdmf_gen_stress.argtypes = [C.c_char, C.c_uint64, C.POINTER(CDataset)]
        lib.cdmf_load_dir.argtypes = [C.c_char_p, C.POINTER(CDataset)]
        lib.cdmf_write_dir.argtypes = [C.c_char_p, C.POINTER(CDataset)]
        lib.cdmf_dataset_free.argtypes = [C.POINTER(CDataset)]
        lib.cdmf_dataset_free.restype = None
        lib.cdmf_data_last_error.restype = C.c_char_p
        lib.cdmf_initial_col.argtypes = [f32p, C.c_uint32, C.c_uint32]
        lib.cdmf_initial_col.restype = None
        lib.cdmf_init_uniform.argtypes = [f32p, C.c_uint64, C.c_float, C.c_uint64]
        lib.cdmf_init_uniform.restype = None
        _data = lib
    return _data


def ptr(a: np.ndarray, ctype):
    return a.ctypes.data_as(C.POINTER(ctype))


def dataset_from_c(cd: CDataset):
    """Copy a library-owned cdmf_dataset into numpy arrays and free it."""
    from .datadir import Dataset

    def take(p, count, dtype):
        if count == 0:
            return np.zeros(0, dtype=dtype)
        return np.ctypeslib.as_array(p, shape=(count,)).copy().astype(dtype, copy=False)

    m, n, nnz, nt = cd.m, cd.n, cd.nnz, cd.nnz_test
    ds = Dataset(m, n,
                 take(cd.csr_ptr, m + 1, np.uint32), take(cd.csr_idx, nnz, np.uint32), take(cd.csr_val, nnz, np.float32),
                 take(cd.csc_ptr, n + 1, np.uint32), take(cd.csc_idx, nnz, np.uint32), take(cd.csc_val, nnz, np.float32),
                 take(cd.test_row, nt, np.uint32), take(cd.test_col, nt, np.uint32), take(cd.test_val, nt, np.float32))
    data_lib().cdmf_dataset_free(C.byref(cd))
    return ds


def dataset_to_c(ds) -> CDataset:
    """Borrowing view of a numpy Dataset as a cdmf_dataset (arrays must stay alive)."""
    cd = CDataset()
    cd.m, cd.n, cd.nnz, cd.nnz_test = ds.m, ds.n, ds.nnz, ds.nnz_test
    cd.max_row_nnz = int(np.diff(ds.csr_ptr.astype(np.int64)).max(initial=0))
    cd.max_col_nnz = int(np.diff(ds.csc_ptr.astype(np.int64)).max(initial=0))
    for name, ct in (("csr_ptr", C.c_uint32), ("csr_idx", C.c_uint32), ("csr_val", C.c_float),
                     ("csc_ptr", C.c_uint32), ("csc_idx", C.c_uint32), ("csc_val", C.c_float),
                     ("test_row", C.c_uint32), ("test_col", C.c_uint32), ("test_val", C.c_float)):
        setattr(cd, name, ptr(getattr(ds, name), ct))
    return cd


def check_data(rc: int) -> None:
    if rc != 0:
        raise RuntimeError("cdmf data layer error %d: %s" % (rc, data_lib().cdmf_data_last_error().decode()))
