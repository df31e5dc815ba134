"""Reach of the matching kernels' loads (tests/cpp/geometry_bounds.cpp, built from the library's own
address functions in csrc/sbm_reach.h): on every workload's geometry -- 480x640, 1024^2, 2048^2, the mixed
batch sizes, the largest image each pyramid accepts; T {4, 8}, {5, 8}, {2, 4, 8}, {8}, {16}; 8 and 16
orientations -- the furthest linear-memory byte of every coarse gather shape and of the refinement stays
inside the zeroed guard, the nonzero-response scatter's row reads stay inside the sentinel padding, its
decode sends the sentinel to the spare word, and no 32-bit offset wraps.  Small levels are checked for
every feature offset.  A guard without the tile overshoot and a padding shorter than one row must fail."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "cpp", "geometry_bounds.cpp")
CSRC = os.path.join(ROOT, "shape_based_matching_amd", "csrc")


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    cxx = shutil.which("g++") or shutil.which("c++")
    if cxx is None or not os.path.isdir("/opt/rocm/include"):
        pytest.skip("no host C++ compiler or HIP headers")
    exe = str(tmp_path_factory.mktemp("bounds") / "geometry_bounds")
    subprocess.check_call([cxx, "-std=c++17", "-O2", "-Wall", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
                           "-I", CSRC, SRC, "-o", exe])
    return exe


def test_every_geometry_stays_inside_guard_and_padding(checker):
    r = subprocess.run([checker], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout[-4000:]
    out = r.stdout
    assert out.strip().endswith(" 0 violations"), out[-2000:]
    for size in ("480x640", "1024x1024", "2048x2048", "960x1280", "1936x2592"):
        assert f"T{{4,8}} {size} level 1 T8 planes 8: coarse loads reach" in out, size
        assert f"T{{4,8}} {size} level 0 T4 planes 8: refine loads reach" in out, size
    assert "T{4,8} largest" in out and "T{2,4,8} 2048x2048 level 2" in out
    assert "exhaustive 32x48 T4 planes 8" in out


@pytest.mark.parametrize("plant", ["guard", "pad"])
def test_planted_short_guard_or_padding_is_caught(checker, plant):
    r = subprocess.run([checker, "--plant", plant], capture_output=True, text=True, timeout=120)
    assert r.returncode == 1 and "VIOLATION" in r.stdout, r.stdout[-2000:]
