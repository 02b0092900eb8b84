"""Build recipes for the native parts of cdmf_amd (run by __graft_entry__.build()).

Outputs (in-tree, git-ignored, shipped with the tree):
  cdmf_amd/_lib/libcdmf_data.so   host data layer (g++, OpenMP)
  cdmf_amd/_lib/libcdmf_hip.so    HIP kernels + C ABI for gfx950 (hipcc)
  cdmf_amd/bin/cdmf               the command-line driver (flags of src/main.cpp)
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
INC = os.path.join(ROOT, "include")
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "_lib")
BIN = os.path.join(HERE, "bin")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

# -packed-fp32-ops: no v_pk_*_f32 (some forms have misbehaved on this pool).
# -ffp-contract=off: residual updates round like the reference (mul, then add).
HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "--offload-arch=" + ARCH,
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-ffp-contract=off", "-munsafe-fp-atomics", "-Wall", "-Wno-unused-result",
]

HIP_SOURCES = [
    "kernels/ccd_kernels.hip",
    "kernels/tiled_kernels.hip",
    "kernels/batch_kernels.hip",
    "kernels/foldin_kernels.hip",
    "kernels/mlp_kernels.hip",
    "kernels/mlp_mfma.hip",
    "ccd_abi.hip",
    "batch.hip",
    "foldin.hip",
]
# plain C++ parts of the HIP library (g++)
LIB_CXX_SOURCES = [
    "tile_layout.cpp",
]
HOST_SOURCES = [
    "host/tools.cpp",
    "host/cdmf_hip.cpp",
    "host/main.cpp",
]


def _run(cmd, cwd=None):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True, cwd=cwd)


def build_data() -> str:
    os.makedirs(LIB, exist_ok=True)
    out = os.path.join(LIB, "libcdmf_data.so")
    _run(["g++", "-O3", "-std=c++17", "-fopenmp", "-fPIC", "-shared", "-Wall", "-I" + INC,
          os.path.join(CSRC, "data", "data.cpp"), "-o", out + ".tmp"])
    os.replace(out + ".tmp", out)
    return out


def build_hip() -> str:
    """One hipcc per source, in parallel (objects under _build/, next to the
    library), then one link."""
    os.makedirs(LIB, exist_ok=True)
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    out = os.path.join(LIB, "libcdmf_hip.so")
    inc = ["-I" + INC, "-I" + CSRC, "-I" + os.path.join(CSRC, "kernels")]
    jobs, objs = [], []
    for s in LIB_CXX_SOURCES:
        obj = os.path.join(objdir, s.replace("/", "_").replace(".cpp", ".o"))
        objs.append(obj)
        cmd = ["g++", "-O2", "-std=c++17", "-fPIC", "-Wall", *inc, "-c", os.path.join(CSRC, s), "-o", obj]
        print("+", " ".join(cmd), flush=True)
        jobs.append((s, subprocess.Popen(cmd)))
    for s in HIP_SOURCES:
        obj = os.path.join(objdir, s.replace("/", "_").replace(".hip", ".o"))
        objs.append(obj)
        cmd = [HIPCC, *HIP_FLAGS, *inc, "-c", os.path.join(CSRC, s), "-o", obj]
        print("+", " ".join(cmd), flush=True)
        jobs.append((s, subprocess.Popen(cmd)))
    failed = [s for s, p in jobs if p.wait() != 0]
    if failed:
        raise RuntimeError("hipcc failed on " + ", ".join(failed))
    _run([HIPCC, *HIP_FLAGS, "-shared", *objs, "-o", out + ".tmp"])
    os.replace(out + ".tmp", out)
    return out


def build_cli() -> str:
    os.makedirs(BIN, exist_ok=True)
    out = os.path.join(BIN, "cdmf")
    srcs = [os.path.join(CSRC, s) for s in HOST_SOURCES]
    _run(["g++", "-O2", "-std=c++17", "-fopenmp", "-Wall", "-I" + INC, "-I" + os.path.join(CSRC, "host"), *srcs,
          "-L" + LIB, "-Wl,-rpath,$ORIGIN/../_lib", "-lcdmf_hip", "-lcdmf_data", "-o", out + ".tmp"])
    os.replace(out + ".tmp", out)
    return out


def build_oracle() -> None:
    if shutil.which("make") is None:
        raise RuntimeError("make not found")
    # the reference-side targets are skipped by the Makefile where the
    # reference checkout is absent
    _run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "all", "refmain"])


def build_all() -> None:
    build_data()
    build_hip()
    build_cli()
    build_oracle()


if __name__ == "__main__":
    what = sys.argv[1:] or ["all"]
    for w in what:
        {"all": build_all, "data": build_data, "hip": build_hip, "cli": build_cli, "oracle": build_oracle}[w]()
