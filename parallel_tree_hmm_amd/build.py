"""Build recipe for libtree_hmm_amd.so (gfx950) and the CPU oracle.

Run as ``python -m parallel_tree_hmm_amd.build`` or through
``__graft_entry__.build()``. Every call recompiles: hipcc cross-compiles the
HIP sources for gfx950 without a GPU present.
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PACKAGE = Path(__file__).resolve().parent
ROOT = PACKAGE.parent
CSRC = PACKAGE / "csrc"
INCLUDE = ROOT / "include"
OBJ = PACKAGE / "_build"
LIBRARY = PACKAGE / "libtree_hmm_amd.so"
ORACLE_DIR = ROOT / "oracle"
ORACLE_LIBRARY = ORACLE_DIR / "liboracle.so"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("THMM_OFFLOAD_ARCH", "gfx950")

# Packed FP32 VALU ops are disabled for every HIP translation unit.
HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Wall",
    "-Xclang",
    "-target-feature",
    "-Xclang",
    "-packed-fp32-ops",
    f"-I{INCLUDE}",
    f"-I{CSRC}",
]
HOST_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-Wall", f"-I{INCLUDE}", f"-I{CSRC}"]

SOURCES = ["plan.cc", "expm.cc", "workspace.cc", "capi.cc", "kernels.hip", "fast_kernels.hip", "static_scratch.hip",
           "device_plan.hip"]


def _run(cmd: list[str]) -> None:
    result = subprocess.run(cmd, capture_output=True, text=True)
    noise = "packed-fp32-ops"
    stderr = "\n".join(
        line for line in result.stderr.splitlines() if noise not in line
    )
    if result.returncode != 0:
        raise RuntimeError(f"command failed: {' '.join(cmd)}\n{stderr}")
    if stderr.strip():
        sys.stderr.write(stderr + "\n")


def _compile(source: str) -> Path:
    src = CSRC / source
    obj = OBJ / (source + ".o")
    if source.endswith(".hip"):
        cmd = [HIPCC, *HIP_FLAGS, "-x", "hip", "-c", str(src), "-o", str(obj)]
    else:
        cmd = [HIPCC, *HOST_FLAGS, "-D__HIP_PLATFORM_AMD__", "-c", str(src), "-o", str(obj)]
    _run(cmd)
    return obj


def build_library() -> Path:
    OBJ.mkdir(exist_ok=True)
    with ThreadPoolExecutor(max_workers=min(4, len(SOURCES))) as pool:
        objects = list(pool.map(_compile, SOURCES))
    tmp = LIBRARY.with_suffix(".so.tmp")
    _run(
        [
            HIPCC,
            f"--offload-arch={ARCH}",
            "-shared",
            "-fPIC",
            *map(str, objects),
            "-Wl,-rpath,/opt/rocm/lib",
            "-o",
            str(tmp),
        ]
    )
    os.replace(tmp, LIBRARY)
    return LIBRARY


def build_oracle() -> Path:
    """The oracle is test infrastructure; it is built beside the product."""
    _run(["make", "-s", "-C", str(ORACLE_DIR)])
    return ORACLE_LIBRARY


CONFORMANCE_SOURCE = ROOT / "tests" / "cpp" / "conformance.cc"
CONFORMANCE_BINARIES = {
    "f64": ROOT / "tests" / "cpp" / "conformance_f64",
    "f32": ROOT / "tests" / "cpp" / "conformance_f32",
}


def build_conformance() -> None:
    """The C++ drop-in conformance programs (tests/cpp), linked against the
    library and the oracle."""
    for precision, binary in CONFORMANCE_BINARIES.items():
        _run(
            [
                "g++",
                "-std=c++20",
                "-O1",
                "-Wall",
                *(["-DTREE_HMM_USE_FLOAT"] if precision == "f32" else []),
                f"-I{INCLUDE}",
                str(CONFORMANCE_SOURCE),
                "-o",
                str(binary),
                f"-L{PACKAGE}",
                "-ltree_hmm_amd",
                f"-L{ORACLE_DIR}",
                "-loracle",
                f"-Wl,-rpath,{PACKAGE}:{ORACLE_DIR}:/opt/rocm/lib",
                "-Wl,-rpath,$ORIGIN/../../parallel_tree_hmm_amd:$ORIGIN/../../oracle",
            ]
        )


def build_all() -> None:
    build_library()
    build_oracle()
    build_conformance()
    shutil.rmtree(OBJ, ignore_errors=True)


if __name__ == "__main__":
    build_all()
    print(f"built {LIBRARY} and {ORACLE_LIBRARY}")
