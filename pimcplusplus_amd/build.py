"""Builds the HIP action library (gfx950) in-tree: pimcplusplus_amd/_lib/libpimc_gpu.so.

    python -m pimcplusplus_amd.build

Every .hip file under csrc/ is compiled to an object with hipcc, in parallel,
then linked into one shared library.  Packed FP32 instructions are disabled
(`-packed-fp32-ops`); contraction follows `#pragma clang fp contract` where the
source sets it.
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "_lib")
OBJDIR = os.path.join(LIBDIR, "_build")
LIB = os.path.join(LIBDIR, "libpimc_gpu.so")
CALIB_SRC = os.path.join(CSRC, "calib", "pimc_calib.hip")
CALIB_LIB = os.path.join(LIBDIR, "libpimc_calib.so")
ARCH = os.environ.get("PIMC_OFFLOAD_ARCH", "gfx950")

HIPCC_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-ffp-contract=fast-honor-pragmas",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result", "-Wno-unused-variable",
]


def hipcc() -> str:
    path = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(path):
        raise RuntimeError("hipcc not found")
    return path


# Per-file flags.  The resident bisection kernel keeps its exp/log constants out
# of loop preheaders (machine LICM would hoist them into registers that stay live
# across the whole kernel and cap the occupancy).
FILE_FLAGS = {
    "pimc_bisection": ["-mllvm", "-disable-machine-licm"],   # every pimc_bisection*.hip
}


def file_flags(src: str) -> list[str]:
    name = os.path.basename(src)
    return [f for prefix, flags in FILE_FLAGS.items() if name.startswith(prefix) for f in flags]


def sources() -> list[str]:
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))


def _up_to_date(obj: str, src: str) -> bool:
    """The object is newer than its source and every header under csrc/ and include/."""
    if not os.path.exists(obj):
        return False
    t = os.path.getmtime(obj)
    deps = [src] + [os.path.join(d, f) for d in (CSRC, os.path.join(HERE, "..", "include"))
                    for f in os.listdir(d) if f.endswith((".hpp", ".h"))]
    return all(os.path.getmtime(d) < t for d in deps)


def _compile(src: str, incremental: bool = False) -> str:
    obj = os.path.join(OBJDIR, os.path.basename(src)[:-4] + ".o")
    if incremental and _up_to_date(obj, src):
        return obj
    cmd = [hipcc(), *HIPCC_FLAGS, *file_flags(src), "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}:\n{r.stderr}")
    return obj


def _link(lib: str, objs: list[str]) -> None:
    tmp = lib + ".tmp"
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, lib)


def build(verbose: bool = False, incremental: bool = False) -> str:
    """Builds libpimc_gpu.so (the action library) and libpimc_calib.so (the benchmark's
    HBM and LDS microbenchmarks, csrc/calib/).  incremental: objects newer than their
    source and the headers are kept (the default recompiles everything)."""
    os.makedirs(OBJDIR, exist_ok=True)
    srcs = sources()
    with ThreadPoolExecutor(max_workers=min(8, len(srcs) + 1)) as ex:
        calib = ex.submit(_compile, CALIB_SRC, incremental)
        objs = list(ex.map(lambda f: _compile(f, incremental), srcs))
        calib_obj = calib.result()
    _link(LIB, objs)
    _link(CALIB_LIB, [calib_obj])
    if verbose:
        print(f"built {LIB} and {CALIB_LIB}")
    return LIB


if __name__ == "__main__":
    build(verbose=True, incremental="--incremental" in sys.argv)
    sys.exit(0)
