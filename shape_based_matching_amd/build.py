"""Build recipe for libsbm_hip.so (gfx950) -- run by __graft_entry__.build() and by hand.

    python -m shape_based_matching_amd.build [--sanitize]

Compiles every source under csrc/ with hipcc for gfx950 into object files under
shape_based_matching_amd/_build/, then links the shared library in-tree next to this file.
"""
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libsbm_hip.so")
INCLUDE = os.path.join(os.path.dirname(HERE), "include")
SOURCES = ["sbm_kernels.hip", "sbm_warp.hip", "sbm_finalize.hip", "sbm_sparse.hip", "sbm_noalloc.hip", "sbm_pipeline.cpp", "sbm_detector.cpp",
           "sbm_train.cpp", "sbm_shapes.cpp", "sbm_yaml.cpp", "sbm_bankfile.cpp", "sbm_server.cpp", "sbm_capi.cpp"]

ARCH = os.environ.get("SBM_OFFLOAD_ARCH", "gfx950")
COMMON = [
    "-O3", "-std=c++17", "-fPIC", "-fvisibility=hidden",
    f"--offload-arch={ARCH}",
    "-ffp-contract=off",                      # the phase step rounds operation by operation
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-result",
    f"-I{INCLUDE}", f"-I{CSRC}",
]


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


# Host code only: AddressSanitizer and UndefinedBehaviorSanitizer on the host pass of every source (the device
# code is built as usual).  The library is then loaded into a process that preloads clang's ASan runtime
# (which carries the UBSan handlers); see tests/test_sanitizers.py.
SANITIZE_HOST = ("-Xarch_host -fsanitize=address -Xarch_host -fsanitize=undefined "
                 "-Xarch_host -fno-sanitize-recover=undefined -Xarch_host -fno-omit-frame-pointer -g").split()
ASAN_DIR = os.path.join(BUILD, "asan")
ASAN_LIB = os.path.join(ASAN_DIR, "libsbm_hip_asan.so")


def _compile(src, sanitize=False):
    out_dir = ASAN_DIR if sanitize else BUILD
    obj = os.path.join(out_dir, src + ".o")
    cmd = [hipcc()] + COMMON + (SANITIZE_HOST if sanitize else []) + ["-c", os.path.join(CSRC, src), "-o", obj]
    if src.endswith(".cpp"):
        cmd += ["-x", "hip"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr}")
    return obj


def build(verbose=False, sanitize=False):
    """Compile and link libsbm_hip.so; sanitize: the host-sanitized copy at _build/asan/libsbm_hip_asan.so."""
    out_dir = ASAN_DIR if sanitize else BUILD
    lib = ASAN_LIB if sanitize else LIB
    os.makedirs(out_dir, exist_ok=True)
    with ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        objs = list(ex.map(lambda src: _compile(src, sanitize), SOURCES))
    tmp = lib + ".tmp"
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs + ["-lz"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, lib)
    if verbose:
        print("built", lib)
    return lib


if __name__ == "__main__":
    build(verbose=True, sanitize="--sanitize" in sys.argv)
    sys.exit(0)
