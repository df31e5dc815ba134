"""The C++ drop-in header (include/line2Dup.h) compiles against the C ABI and, on the GPU, returns the
same match list as the Python binding and the oracle."""
import os
import shutil
import subprocess

import numpy as np
import pytest

import oracle
import shape_based_matching_amd as sbm
from helpers import FIXTURE_BANKS, bank_of, golden, load_fixture_detector, scene_bgr, to_gray

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "cpp", "drop_in_match.cpp")


def build(tmp_path):
    cxx = shutil.which("g++") or shutil.which("c++")
    if cxx is None:
        pytest.skip("no host C++ compiler")
    exe = str(tmp_path / "drop_in_match")
    libdir = os.path.dirname(sbm.LIB_PATH)
    cmd = [cxx, "-std=c++17", "-O2", "-Wall", "-I", os.path.join(ROOT, "include"), SRC, "-o", exe,
           "-L", libdir, "-lsbm_hip", f"-Wl,-rpath,{libdir}"]
    subprocess.check_call(cmd)
    return exe


def test_cpp_header_compiles_and_links(tmp_path):
    assert os.path.exists(build(tmp_path))


@pytest.mark.gpu
def test_cpp_caller_matches_python_and_oracle(tmp_path):
    exe = build(tmp_path)
    img = to_gray(scene_bgr(2))
    raw = tmp_path / "img.raw"
    raw.write_bytes(img.tobytes())
    out = subprocess.run([exe, golden(FIXTURE_BANKS[2][0]), str(raw), str(img.shape[0]), str(img.shape[1]), "1", "90"],
                         capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    got = []
    for line in out.stdout.split("\n"):
        if line.strip():
            x, y, bits, c, t = line.split()
            got.append((int(x), int(y), np.uint32(int(bits, 16)).view(np.float32), c, int(t)))
    det = load_fixture_detector(2)
    py = [(m.x, m.y, np.float32(m.similarity), m.class_id, m.template_id) for m in det.match(img, 90)]
    ref = [(x, y, np.float32(s), c, t) for (x, y, s, c, t) in oracle.match(img, [4, 8], 90, bank_of(det))]
    assert got == py == ref and len(ref) > 0
