"""CPU sanitizer runs of the host-side C++ and the oracle (no GPU).

tests/sanitize/host_driver.cpp drives graph loading and validation (good and
malformed files), the edge-list converter, the generators, orientation, the
out-of-core host work (row blocks, block copies, flag merges), the result
writer, and the oracle's OpenMP loops (whole pass, phases after similarity,
prep counts, threaded per-edge similarity), checking them against each
other.  It is built twice with ROCm's clang:

- ThreadSanitizer with LLVM's OpenMP runtime and its Archer tool, which
  gives the sanitizer OpenMP's own synchronisation (without it, every
  barrier inside the runtime reads as a race); 4 threads.  A planted race in
  the driver must be reported, so a clean run means something.
- AddressSanitizer + UndefinedBehaviorSanitizer; a planted heap read out of
  bounds must be reported.
"""
from __future__ import annotations

import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "efficientstructuralgraphclusteringongpus_amd", "csrc")
LLVM = "/opt/rocm/llvm"
CLANG = os.path.join(LLVM, "bin", "clang")
CLANGXX = os.path.join(LLVM, "bin", "clang++")
ARCHER = os.path.join(LLVM, "lib", "libarcher.so")
ROCM_INCLUDE = "/opt/rocm/include"

pytestmark = pytest.mark.skipif(not (os.path.exists(CLANGXX) and os.path.exists(ARCHER)),
                                reason="ROCm clang with libarcher.so is needed for the sanitizer builds")


def build(tmp, name, flags):
    """Compiles the driver, the three host files and the oracle with `flags`."""
    out = os.path.join(tmp, name)
    objs = []
    common = ["-g", "-O1", "-fopenmp", "-fno-omit-frame-pointer", *flags]
    inc = ["-I" + os.path.join(ROOT, "include"), "-I" + CSRC, "-I" + os.path.join(ROOT, "oracle"),
           "-I" + ROCM_INCLUDE, "-D__HIP_PLATFORM_AMD__"]
    for src in (os.path.join(ROOT, "tests", "sanitize", "host_driver.cpp"), os.path.join(CSRC, "graph_host.cpp"),
                os.path.join(CSRC, "ooc_host.cpp"), os.path.join(CSRC, "result_io.cpp")):
        obj = os.path.join(tmp, name + "_" + os.path.basename(src) + ".o")
        subprocess.run([CLANGXX, "-std=c++17", *common, *inc, "-c", src, "-o", obj], check=True)
        objs.append(obj)
    obj = os.path.join(tmp, name + "_scan_oracle.o")
    subprocess.run([CLANG, "-std=c11", "-ffp-contract=off", *common, "-I" + os.path.join(ROOT, "oracle"), "-c",
                    os.path.join(ROOT, "oracle", "scan_oracle.c"), "-o", obj], check=True)
    objs.append(obj)
    subprocess.run([CLANGXX, *common, *objs, "-o", out, "-Wl,-rpath," + os.path.join(LLVM, "lib")], check=True)
    return out


def run(exe, args, env_extra, tmp):
    env = dict(os.environ, OMP_NUM_THREADS="4", **env_extra)
    return subprocess.run([exe, *args], env=env, cwd=tmp, capture_output=True, text=True, timeout=600)


@pytest.fixture(scope="module")
def scratch(tmp_path_factory):
    d = tmp_path_factory.mktemp("sanitize")
    yield str(d)
    shutil.rmtree(d, ignore_errors=True)


TSAN_ENV = {"TSAN_OPTIONS": "halt_on_error=0,ignore_noninstrumented_modules=1,exitcode=66",
            "OMP_TOOL_LIBRARIES": ARCHER}


def test_thread_sanitizer_clean(scratch):
    exe = build(scratch, "tsan", ["-fsanitize=thread"])
    work = os.path.join(scratch, "tsan_run")
    os.makedirs(work, exist_ok=True)
    r = run(exe, ["run", work], TSAN_ENV, scratch)
    assert "Archer detected OpenMP application with TSan" in r.stderr or "ThreadSanitizer" not in r.stderr
    assert "ThreadSanitizer" not in r.stderr, r.stderr[-4000:]
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert "0 failed checks" in r.stdout
    # the planted race is reported: the tool is live in this instrumented binary
    p = run(exe, ["race"], TSAN_ENV, scratch)
    assert "ThreadSanitizer: data race" in p.stderr and p.returncode == 66


def test_address_and_undefined_sanitizers_clean(scratch):
    exe = build(scratch, "asan", ["-fsanitize=address,undefined", "-fno-sanitize-recover=undefined"])
    work = os.path.join(scratch, "asan_run")
    os.makedirs(work, exist_ok=True)
    # LLVM's OpenMP runtime keeps one allocation of its own until exit
    supp = os.path.join(scratch, "lsan.supp")
    with open(supp, "w") as f:
        f.write("leak:libomp.so\n")
    r = run(exe, ["run", work], {"ASAN_OPTIONS": "detect_leaks=1", "LSAN_OPTIONS": "suppressions=" + supp}, scratch)
    assert "AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr, r.stderr[-4000:]
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert "0 failed checks" in r.stdout
    # the planted out-of-bounds read is reported: the tool is live in this instrumented binary
    p = run(exe, ["overflow"], {}, scratch)
    assert "heap-buffer-overflow" in p.stderr and p.returncode != 0
