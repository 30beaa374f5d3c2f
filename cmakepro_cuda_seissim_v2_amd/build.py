"""Build recipes for the native libraries (hipcc for gfx950, g++ for the oracle).

``build_library()`` compiles the HIP solver into ``cmakepro_cuda_seissim_v2_amd/lib/libseis3d.so``;
``build_oracle()`` compiles the CPU oracle (test infrastructure) into ``oracle/lib/liboracle3d.so``.
Both build in-tree so the shared objects travel with the source tree.
"""
from __future__ import annotations

import os
import subprocess
from pathlib import Path

PACKAGE_DIR = Path(__file__).resolve().parent
REPO_DIR = PACKAGE_DIR.parent
CSRC = PACKAGE_DIR / "csrc"
LIB_DIR = PACKAGE_DIR / "lib"
ORACLE_DIR = REPO_DIR / "oracle"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX", "g++")

# -ffp-contract=off: no fused multiply-add, so the kernels round every
# operation exactly like the FP32 CPU oracle.  Packed FP32 is disabled.
HIP_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]

LIBRARIES = {
    "libseis2d.so": ["seis2d_kernels.hip", "seis2d_api.cpp", "seis2d_io.cpp"],
    "libseis3d.so": ["seis3d_kernels.hip", "seis3d_dma.hip", "seis3d_adjoint.hip",
                    "seis3d_abc.hip", "seis3d_stream.hip", "seis3d_noalloc.hip", "seis3d_api.cpp", "seis3d_fwi.cpp", "seis3d_abc.cpp", "seis3d_io.cpp"],
}


def _run(cmd: list[str]) -> None:
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if proc.returncode != 0:
        raise RuntimeError(f"build failed ({proc.returncode}): {' '.join(cmd)}\n{proc.stdout}")


def build_library(name: str = "libseis3d.so") -> Path:
    """Compile one HIP shared library for gfx950 (always rebuilds)."""
    LIB_DIR.mkdir(parents=True, exist_ok=True)
    out = LIB_DIR / name
    objs = []
    for src in LIBRARIES[name]:
        obj = LIB_DIR / (src.replace(".", "_") + ".o")
        lang = ["-x", "hip"] if src.endswith(".hip") else []
        _run([HIPCC, *HIP_FLAGS, "-c", *lang, str(CSRC / src), "-o", str(obj)])
        objs.append(str(obj))
    tmp = out.with_suffix(".so.tmp")
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-o", str(tmp)])
    os.replace(tmp, out)
    for o in objs:
        Path(o).unlink(missing_ok=True)
    return out


def build_oracle() -> Path:
    """Compile the CPU oracle (checker only) with OpenMP and no FP contraction."""
    out_dir = ORACLE_DIR / "lib"
    out_dir.mkdir(parents=True, exist_ok=True)
    out = out_dir / "liboracle3d.so"
    tmp = out.with_suffix(".so.tmp")
    _run([CXX, "-O3", "-march=x86-64-v3", "-ffp-contract=off", "-fopenmp", "-std=c++17", "-fPIC", "-shared",
          "-Wall", str(ORACLE_DIR / "oracle_capi.cpp"), str(ORACLE_DIR / "oracle2d_capi.cpp"), "-o", str(tmp)])
    os.replace(tmp, out)
    return out


CPP_TESTS = {"dropin_session": ["seis3d", "seis2d"], "seis3d_io_regression": ["seis3d"],
             "seis2d_record_regression": ["seis2d"]}


def build_cpp_tests() -> list[Path]:
    """Compile the C++ test programs (tests/cpp) against include/Seis3D.hpp / Seis2D.hpp and their libraries."""
    src_dir = REPO_DIR / "tests" / "cpp"
    out_dir = src_dir / "bin"
    out_dir.mkdir(parents=True, exist_ok=True)
    outs = []
    for name, libs in CPP_TESTS.items():
        out = out_dir / name
        _run([CXX, "-O2", "-std=c++17", "-Wall", f"-I{REPO_DIR / 'include'}", str(src_dir / f"{name}.cpp"),
              "-o", str(out), f"-L{LIB_DIR}", *[f"-l{lib}" for lib in libs],
              f"-Wl,-rpath,$ORIGIN/../../../cmakepro_cuda_seissim_v2_amd/lib"])
        outs.append(out)
    return outs


def build_all() -> None:
    for name in LIBRARIES:
        build_library(name)
    build_oracle()
    build_cpp_tests()


if __name__ == "__main__":
    build_all()
    print("built", ", ".join(LIBRARIES), "and the oracle")
