"""The CPU suite again, under AddressSanitizer and UndefinedBehaviorSanitizer.

Both native libraries are rebuilt with the sanitizers on their host code: libsbm_hip (hipcc, the
`-Xarch_host` sanitizer flags, device code as usual: the host-side C ABI, YAML/bank parsing, argument
checks, result decoding) and the oracle with the vectorised CPU baseline (clang). A child pytest process
preloads clang's ASan runtime, points the loaders at the sanitized builds (SBM_HIP_LIBRARY,
SBM_ORACLE_LIBRARY) and runs every CPU test but the C++ drop-in (which links the unsanitized library into
a program of its own) and this file; any report ends it with a non-zero status.
"""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLANG = "/opt/rocm/llvm/bin/clang"

pytestmark = pytest.mark.skipif(os.environ.get("SBM_UNDER_SANITIZERS") == "1" or not os.path.exists(CLANG),
                                reason="already sanitized / no clang")


def _asan_runtime():
    r = subprocess.run([CLANG, "--print-file-name=libclang_rt.asan-x86_64.so"], capture_output=True, text=True)
    path = r.stdout.strip()
    if r.returncode != 0 or not os.path.isabs(path) or not os.path.exists(path):
        pytest.skip("clang's ASan runtime is not installed")
    return path


@pytest.fixture(scope="module")
def sanitized_env():
    from shape_based_matching_amd import build as hbuild
    rt = _asan_runtime()
    hip_lib = hbuild.build(sanitize=True)
    r = subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "asan"], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    env = dict(os.environ)
    env.update(LD_PRELOAD=rt, SBM_UNDER_SANITIZERS="1",
               ASAN_OPTIONS="detect_leaks=0:halt_on_error=1:abort_on_error=0",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1",
               SBM_HIP_LIBRARY=hip_lib,
               SBM_ORACLE_LIBRARY=os.path.join(ROOT, "oracle", "_asan", "liboracle_asan.so"),
               LD_LIBRARY_PATH=os.pathsep.join(p for p in ("/opt/rocm/llvm/lib", os.environ.get("LD_LIBRARY_PATH"))
                                               if p))
    return env


def test_sanitized_libraries_are_the_ones_loaded(sanitized_env):
    probe = ("import oracle, shape_based_matching_amd._lib as L; L.lib(); oracle.py_oracle.lib(); "
             "m = open('/proc/self/maps').read(); "
             "print(int('libsbm_hip_asan.so' in m), int('liboracle_asan.so' in m), int('libclang_rt.asan' in m))")
    r = subprocess.run([sys.executable, "-c", probe], cwd=ROOT, env=sanitized_env, capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stderr[-3000:]
    assert r.stdout.split() == ["1", "1", "1"], r.stdout


def test_cpu_suite_under_sanitizers(sanitized_env):
    cmd = [sys.executable, "-m", "pytest", "tests", "-x", "-q", "-m", "not gpu", "-p", "no:cacheprovider",
           "--deselect", "tests/test_cpp_dropin.py", "--deselect", "tests/test_sanitizers.py"]
    r = subprocess.run(cmd, cwd=ROOT, env=sanitized_env, capture_output=True, text=True, timeout=1200)
    out = r.stdout[-4000:] + r.stderr[-6000:]
    assert r.returncode == 0, out
    assert " passed" in r.stdout and "error" not in r.stdout.splitlines()[-1], out
