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
TRAIN_SRC = os.path.join(ROOT, "tests", "cpp", "drop_in_train.cpp")


def build(tmp_path, src=SRC, defines=()):
    cxx = shutil.which("g++") or shutil.which("c++")
    if cxx is None:
        pytest.skip("no host C++ compiler")
    exe = str(tmp_path / (os.path.splitext(os.path.basename(src))[0] + "".join("_" + d.lower() for d in defines)))
    libdir = os.path.dirname(sbm.LIB_PATH)
    cmd = [cxx, "-std=c++17", "-O2", "-Wall", "-I", os.path.join(ROOT, "include")] + [f"-D{d}" for d in defines] + \
          [src, "-o", exe, "-L", libdir, "-lsbm_hip", f"-Wl,-rpath,{libdir}"]
    subprocess.check_call(cmd)
    return exe


def test_cpp_header_compiles_and_links(tmp_path):
    assert os.path.exists(build(tmp_path))
    assert os.path.exists(build(tmp_path, TRAIN_SRC))
    assert os.path.exists(build(tmp_path, defines=("LINE2DUP_16BIT_ORI",)))


@pytest.mark.gpu
def test_cpp_caller_of_the_16_orientation_variant(tmp_path):
    """The same caller built with LINE2DUP_16BIT_ORI reads a 16-orientation bank and matches a colour scene as
    the Python binding's Detector16 and the 16-orientation oracle do."""
    from PIL import Image
    exe = build(tmp_path, defines=("LINE2DUP_16BIT_ORI",))
    train = np.ascontiguousarray(np.asarray(Image.open(golden("case1_train.png")).convert("RGB"))[:, :, ::-1])
    det = sbm.Detector16(63, [4, 8])
    assert det.add_templates_transformed(train, "test", [(a, 1.0) for a in range(0, 360, 30)]) == list(range(12))
    det.writeClass("test", tmp_path / "bank16.yaml")
    img = np.ascontiguousarray(scene_bgr(1))
    raw = tmp_path / "img.raw"
    raw.write_bytes(img.tobytes())
    out = subprocess.run([exe, str(tmp_path / "bank16.yaml"), str(raw), str(img.shape[0]), str(img.shape[1]), "3",
                          "50"], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    got = []
    for line in out.stdout.split("\n"):
        if line.strip():
            x, y, bits, c, t = line.split()
            got.append((int(x), int(y), np.uint32(int(bits, 16)).view(np.float32), c, int(t)))
    py = [(m.x, m.y, np.float32(m.similarity), m.class_id, m.template_id) for m in det.match(img, 50)]
    ref = [(x, y, np.float32(s), c, t) for (x, y, s, c, t) in oracle.match16(img, [4, 8], 50, bank_of(det))]
    assert got == py == ref and len(ref) > 0


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


@pytest.mark.gpu
def test_cpp_training_loop_equals_python_and_oracle(tmp_path):
    """shapeInfo_producer + addTemplate + writeClasses + save_infos from C++ give the same bank and
    infos as the Python binding's batched training, and the oracle's templates."""
    from PIL import Image
    exe = build(tmp_path, TRAIN_SRC)
    bgr = np.ascontiguousarray(np.asarray(Image.open(golden("case2_train.png")).convert("RGB"))[:, :, ::-1])
    raw = tmp_path / "train.raw"
    raw.write_bytes(bgr.tobytes())
    prefix = str(tmp_path) + "/"
    out = subprocess.run([exe, str(raw), str(bgr.shape[0]), str(bgr.shape[1]), "3", "30", "15", prefix],
                         capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    ids = [int(v) for v in out.stdout.split()]
    assert ids == list(range(25))  # 0..360 step 15
    det = sbm.Detector(30, [4, 8])
    assert det.readClass(prefix + "test_templ.yaml") == "test"
    py = sbm.Detector(30, [4, 8])
    sp = sbm.shapeInfo_producer(bgr)
    sp.angle_range, sp.angle_step = [0, 360], 15
    sp.produce_infos()
    assert py.add_templates_transformed(bgr, "test", sp.infos) == ids
    py.writeClass("test", tmp_path / "py.yaml")
    assert (tmp_path / "py.yaml").read_bytes() == open(prefix + "test_templ.yaml", "rb").read()
    assert sbm.shapeInfo_producer.load_infos(prefix + "test_info.yaml") == sp.infos
    for k, info in enumerate(sp.infos):
        ref = oracle.add_template(oracle.shape_transform(bgr, info.angle, info.scale), 30, 2,
                                  mask=oracle.mask_of(sp.mask, info.angle, info.scale))
        got = det.getTemplates("test", k)
        for lvl in range(2):
            assert np.array_equal(got[lvl].features, ref[lvl]["features"]), (k, lvl)
