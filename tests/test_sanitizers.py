"""Address and undefined-behaviour sanitizer builds of the host code, run under part of the CPU
suite in child processes (the sanitizer runtime is preloaded into a fresh interpreter):

  - the CPU oracle built by g++ with -fsanitize=address,undefined, running the oracle pin tests of
    FormU's boundary rules, the ghost fills, GradP, the Krylov solver and the Runge-Kutta
    predictors with their tangent;
  - the HIP library with its host code built by hipcc with address sanitizing and trapping
    undefined-behaviour checks (device code unchanged), running the C-ABI tests: library load,
    exports, struct layouts and every host-side error path.
A sanitizer report, or a test failure, fails the test.  These tests build into a temporary
directory and never replace the in-tree libraries.
"""
import os
import shutil
import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]


def _run(cmd, env, timeout):
    p = subprocess.run(cmd, cwd=ROOT, env=env, capture_output=True, text=True, timeout=timeout)
    out = p.stdout + p.stderr
    return p.returncode, out


def _gcc_runtime(name):
    p = subprocess.run(["g++", f"-print-file-name={name}"], capture_output=True, text=True)
    path = p.stdout.strip()
    return path if path and Path(path).exists() else None


def test_oracle_under_asan_ubsan(tmp_path):
    if not shutil.which("g++"):
        pytest.skip("g++ not available")
    asan = _gcc_runtime("libasan.so")
    if not asan:
        pytest.skip("libasan not available")
    lib = tmp_path / "liboracle_san.so"
    cmd = ["g++", "-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
           "-fno-sanitize-recover=undefined", "-ffp-contract=off", "-fopenmp", "-fPIC", "-std=c++17", "-shared",
           "-o", str(lib), "oracle/tosca_oracle.cpp", "oracle/tosca_oracle_tangent.cpp"]
    subprocess.run(cmd, cwd=ROOT, check=True, timeout=600)
    env = dict(os.environ, LD_PRELOAD=asan, TOSCA_ORACLE_LIB=str(lib), ASAN_OPTIONS="detect_leaks=0",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1", OMP_NUM_THREADS="2")
    tests = ["tests/test_oracle_boundary_pins.py", "tests/test_oracle_bcs.py", "tests/test_oracle_inflow.py",
             "tests/test_oracle_gradp.py",
             "tests/test_oracle_krylov.py", "tests/test_oracle_rk3w.py", "tests/test_oracle_tangent.py"]
    rc, out = _run([sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider", "-m", "not gpu", *tests],
                   env, 1500)
    tail = out[-3000:]
    assert "ERROR: AddressSanitizer" not in out, tail
    assert "runtime error:" not in out, tail
    assert rc == 0, tail


def test_host_abi_under_asan_ubsan(tmp_path):
    from tosca_amd import build as B
    try:
        cc = B.hipcc()
    except RuntimeError:
        pytest.skip("hipcc not available")
    rt = sorted(Path("/opt/rocm/llvm/lib/clang").glob("*/lib/linux/libclang_rt.asan-x86_64.so"))
    if not rt:
        pytest.skip("clang address-sanitizer runtime not available")
    host_san = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
                "-Xarch_host", "-fsanitize-trap=undefined", "-Xarch_host", "-fno-omit-frame-pointer"]
    objs = []
    for src, extra in B.SOURCES.items():
        obj = tmp_path / (Path(src).stem + ".o")
        cmd = [cc, *B.COMMON, *extra, *host_san, "-c", str(B.CSRC / src), "-o", str(obj)]
        subprocess.run(cmd, check=True, timeout=900, capture_output=True)
        objs.append(str(obj))
    lib = tmp_path / "libtosca_rhs_san.so"
    subprocess.run([cc, f"--offload-arch={B.ARCH}", "-shared", "-fPIC", *objs, "-o", str(lib)], check=True, timeout=600,
                   capture_output=True)
    env = dict(os.environ, LD_PRELOAD=str(rt[-1]), TOSCA_LIB=str(lib), ASAN_OPTIONS="detect_leaks=0")
    rc, out = _run([sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider", "-m", "not gpu",
                    "tests/test_abi_cpu.py"], env, 900)
    tail = out[-3000:]
    assert "ERROR: AddressSanitizer" not in out, tail
    assert rc == 0, tail
