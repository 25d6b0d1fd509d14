"""The oracle (and the product's host generator it is fed by) under the
compiler's sanitizers, as standalone executables built from
oracle/sanitize/: AddressSanitizer + UndefinedBehaviorSanitizer, and
ThreadSanitizer over the oracle's OpenMP event loop (LLVM's OpenMP runtime
with its Archer tool, so that the sanitizer sees OpenMP's barriers).  Each
run reconstructs generated events (reference caps at Poisson(300), lifted
caps at Poisson(1200); full chain, vertexing, dimuons) with one and with four
threads and requires identical results and no sanitizer report."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN = os.path.join(ROOT, "oracle", "sanitize")
GEOMETRY = os.path.join(ROOT, "e1039-gpu_amd", "data", "basic_geometry_e906.txt")
LLVM_LIB = "/opt/rocm/llvm/lib"


@pytest.fixture(scope="module")
def built(tmp_path_factory):
    if not shutil.which("g++"):
        pytest.skip("g++ not found")
    out = tmp_path_factory.mktemp("san")
    targets = [str(out / "oracle_asan")]
    tsan = os.path.exists("/opt/rocm/llvm/bin/clang++") and os.path.exists(os.path.join(LLVM_LIB, "libarcher.so"))
    if tsan:
        targets.append(str(out / "oracle_tsan"))
    subprocess.run(["make", "-s", "-C", SAN, f"OUT={out}", "-j2"] + targets, check=True, timeout=600)
    return out, tsan


def run(exe, env, *args):
    p = subprocess.run([exe, GEOMETRY, *args], capture_output=True, text=True, env={**os.environ, **env},
                       timeout=600)
    return p.returncode, p.stdout, p.stderr


def test_asan_ubsan(built):
    out, _ = built
    rc, so, se = run(str(out / "oracle_asan"), {"ASAN_OPTIONS": "detect_leaks=1"}, "4", "150", "6")
    assert rc == 0, se[-3000:]
    assert "ERROR: AddressSanitizer" not in se and "runtime error" not in se, se[-3000:]
    assert so.count("identical") == 2, so


def test_tsan_openmp(built):
    out, tsan = built
    if not tsan:
        pytest.skip("clang++ with LLVM's OpenMP Archer tool not found")
    env = {"OMP_TOOL_LIBRARIES": os.path.join(LLVM_LIB, "libarcher.so"),
           "TSAN_OPTIONS": "halt_on_error=1 ignore_noninstrumented_modules=1 suppressions="
                           + os.path.join(SAN, "tsan.supp")}
    rc, so, se = run(str(out / "oracle_tsan"), env, "4", "60", "2")
    assert rc == 0, se[-3000:]
    assert "WARNING: ThreadSanitizer" not in se, se[-3000:]
    assert so.count("identical") == 2, so
