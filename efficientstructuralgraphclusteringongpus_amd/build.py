"""Build recipe for libsgc.so and the GPUScan command line (gfx950).

Host-only sources are compiled by g++ with OpenMP (libgomp), HIP sources by
hipcc for gfx950 with FP contraction off (the reference's double
expressions must round exactly as written) and packed FP32 disabled; the
shared library is linked in-tree under ``lib/`` so it travels with the tree.

    python -m efficientstructuralgraphclusteringongpus_amd.build
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIBDIR = os.path.join(PKG, "lib")
OBJDIR = os.path.join(PKG, "_build")
INCLUDE = os.path.join(ROOT, "include")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = shutil.which("hipcc") or os.path.join(ROCM, "bin", "hipcc")
CXX = os.environ.get("CXX", "g++")
ARCH = "gfx950"

HOST_SOURCES = ["graph_host.cpp", "ooc_host.cpp", "result_io.cpp"]
HIP_SOURCES = ["kernels.hip", "similarity.hip", "device_api.hip", "ooc.hip", "serve.hip"]

COMMON = ["-O3", "-fPIC", "-std=c++17", "-I" + INCLUDE, "-I" + CSRC, "-Wall", "-Wno-unused-function"]
HOST_FLAGS = COMMON + ["-fopenmp", "-D__HIP_PLATFORM_AMD__", "-I" + os.path.join(ROCM, "include"),
                       "-ffp-contract=off"]
HIP_FLAGS = COMMON + [
    "--offload-arch=" + ARCH,
    "-ffp-contract=off",
    "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics",
]


def _run(cmd: list[str]) -> None:
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in proc.stdout.splitlines()
                    if "packed-fp32-ops" not in l and "is not a recognized feature" not in l)
    if proc.returncode != 0:
        raise RuntimeError("build step failed:\n  " + " ".join(cmd) + "\n" + out)


# similarity.hip: its atomics are single-lane (the unit queue, claim cursors,
# per-edge counters).  The AMDGPU atomic optimizer would turn each into a wave
# reduction that reads the result at once; without it, the queue's next unit
# is taken while the current one runs.
FILE_FLAGS = {"similarity.hip": ["-mllvm", "-amdgpu-atomic-optimizer-strategy=None"]}


def _extra_defines() -> list[str]:
    """-D flags from SGC_BUILD_DEFINES (comma separated), for debug builds."""
    return ["-D" + d for d in os.environ.get("SGC_BUILD_DEFINES", "").split(",") if d]


def _compile(src: str) -> str:
    obj = os.path.join(OBJDIR, os.path.basename(src) + ".o")
    path = os.path.join(CSRC, src)
    if src.endswith(".hip"):
        _run([HIPCC, *HIP_FLAGS, *FILE_FLAGS.get(src, []), *_extra_defines(), "-c", path, "-o", obj])
    else:
        _run([CXX, *HOST_FLAGS, "-c", path, "-o", obj])
    return obj


def build(verbose: bool = True) -> str:
    """Compiles every source and links lib/libsgc.so and lib/GPUScan."""
    os.makedirs(LIBDIR, exist_ok=True)
    os.makedirs(OBJDIR, exist_ok=True)
    sources = HOST_SOURCES + HIP_SOURCES
    with ThreadPoolExecutor(max_workers=min(8, len(sources))) as ex:
        objs = list(ex.map(_compile, sources))
    lib = os.path.join(LIBDIR, "libsgc.so")
    tmp = lib + ".tmp"
    _run([HIPCC, "-shared", "--offload-arch=" + ARCH, "-o", tmp, *objs, "-lgomp", "-lpthread"])
    os.replace(tmp, lib)
    cli = os.path.join(LIBDIR, "GPUScan")
    _run([CXX, "-O2", "-std=c++17", "-I" + INCLUDE, os.path.join(CSRC, "gpuscan_cli.cpp"), "-o", cli + ".tmp",
          "-L" + LIBDIR, "-lsgc", "-Wl,-rpath,$ORIGIN"])
    os.replace(cli + ".tmp", cli)
    if verbose:
        print(f"built {lib} and {cli}")
    return lib


if __name__ == "__main__":
    try:
        build()
    except RuntimeError as e:
        print(e, file=sys.stderr)
        sys.exit(1)
