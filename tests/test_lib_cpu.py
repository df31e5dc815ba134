"""CPU-side checks of libsbm_hip.so: it loads, exports every declared entry point, and its host-only
parts (bank persistence, validation, NMS) behave as the reference's.  No GPU work is launched."""
import ctypes
import gzip
import os
import time

import numpy as np
import pytest

import oracle
import shape_based_matching_amd as sbm
from helpers import FIXTURE_BANKS, golden
from shape_based_matching_amd import _lib


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    names = _lib.header_symbols()
    assert len(names) >= 40
    for n in names:
        assert hasattr(L, n), n
        assert n in _lib._SIGS, f"{n} has no ctypes signature"


@pytest.mark.parametrize("case", [0, 1, 2])
def test_yaml_round_trip_is_byte_exact(case, tmp_path):
    """readClass then writeClass reproduces the reference's FileStorage output byte for byte."""
    src = golden(FIXTURE_BANKS[case][0])
    raw = gzip.open(src).read()
    det = sbm.Detector(63, [4, 8])
    cid = det.readClass(src)
    assert cid == FIXTURE_BANKS[case][1]
    out = tmp_path / "o.yaml"
    det.writeClass(cid, out)
    assert out.read_bytes() == raw
    # .gz written by writeClasses reads back to the same bank
    det.writeClasses(str(tmp_path / "t_%s.yml.gz"))
    det2 = sbm.Detector(63, [4, 8])
    det2.readClasses([cid], str(tmp_path / "t_%s.yml.gz"))
    assert det2.numTemplates(cid) == det.numTemplates(cid)
    for tid in (0, det.numTemplates(cid) - 1):
        for a, b in zip(det.getTemplates(cid, tid), det2.getTemplates(cid, tid)):
            assert np.array_equal(a.features, b.features) and (a.width, a.tl_x) == (b.width, b.tl_x)


def test_fixture_bank_contents():
    det = sbm.Detector(63, [4, 8])
    det.readClass(golden(FIXTURE_BANKS[0][0]))
    assert det.numTemplates() == 89 and det.classIds() == ["circle"]
    t = det.getTemplates("circle", 0)
    assert (t[0].width, t[0].height, t[0].tl_x, t[0].tl_y, t[0].pyramid_level) == (39, 39, 380, 380, 0)
    assert t[0].features[:3].tolist() == [[11, 2, 3], [29, 2, 5], [2, 11, 1]]
    assert (t[1].width, t[1].tl_x, t[1].pyramid_level) == (19, 190, 1)


def test_read_class_duplicate_and_override(tmp_path):
    det = sbm.Detector(63, [4, 8])
    det.readClass(golden(FIXTURE_BANKS[2][0]))
    with pytest.raises(sbm.SbmError) as e:
        det.readClass(golden(FIXTURE_BANKS[2][0]))
    assert e.value.code == _lib.SBM_EDUPCLASS
    # an override onto an existing class keeps the existing templates (std::map::insert)
    n = det.numTemplates("test")
    det.readClass(golden(FIXTURE_BANKS[0][0]), "test")
    assert det.numTemplates("test") == n
    assert det.readClass(golden(FIXTURE_BANKS[0][0]), "circle2") == "circle2"
    assert det.classIds() == ["circle2", "test"]


def test_detector_write_read(tmp_path):
    det = sbm.Detector(150, [4, 8], 25.0, 55.5)
    p = tmp_path / "d.yaml"
    det.write(p)
    text = p.read_text()
    assert text.startswith("%YAML:1.0\n---\n")
    assert "T: [ 4, 8 ]" in text and "weak_threshold: 25." in text and "strong_threshold: 5.5500000000000000e+01" in text
    d2 = sbm.Detector(63, [2])
    d2.readClass(golden(FIXTURE_BANKS[0][0]))
    d2.read(p)
    # Detector::read clears the bank, restores levels and T, and resets the modality
    assert d2.pyramidLevels() == 2 and d2.getT(1) == 8 and d2.numTemplates() == 0
    assert d2.getModalities() == dict(weak_threshold=30.0, num_features=63, strong_threshold=60.0)


def test_match_validates_sizes_before_touching_the_gpu():
    det = sbm.Detector(63, [4, 8])
    det.readClass(golden(FIXTURE_BANKS[2][0]))
    for shape in [(1000, 1000), (30, 32), (64, 36)]:
        with pytest.raises(sbm.SbmError) as e:
            det.match(np.zeros(shape, np.uint8), 90)
        assert e.value.code == _lib.SBM_ESIZE
    with pytest.raises(sbm.SbmError) as e:
        det.match(np.zeros((64, 64), np.uint8), 90, masks=np.zeros((32, 32), np.uint8))
    with pytest.raises(sbm.SbmError):
        sbm.Detector(63, [0])


def test_detector16_constructors_and_rules_before_touching_the_gpu(tmp_path):
    """Detector(), Detector(T), Detector(num_features, T) of line2Dup_16bit_ori.cpp:1036-1056."""
    d = sbm.Detector16()
    assert d.orientationBins() == 16 and d.pyramidLevels() == 2 and [d.getT(0), d.getT(1)] == [5, 8]
    assert d.getModalities() == dict(weak_threshold=10.0, num_features=63, strong_threshold=55.0)
    d2 = sbm.Detector16(128, [4, 8])
    assert d2.getModalities()["num_features"] == 128 and [d2.getT(0), d2.getT(1)] == [4, 8]
    assert sbm.Detector16([2, 4, 8]).pyramidLevels() == 3
    assert sbm.Detector().orientationBins() == 8
    with pytest.raises(sbm.SbmError) as e:
        d.match(np.zeros((80, 80), np.uint8), 50)  # the variant reads three channels
    assert e.value.code == _lib.SBM_EINVAL
    with pytest.raises(sbm.SbmError) as e:
        d.addTemplate_rotate("a", 0, 10.0, (1.0, 1.0))
    assert e.value.code == _lib.SBM_EINVAL
    # labels up to 15 survive the YAML round trip
    g = np.random.default_rng(0)
    f = [np.column_stack([g.integers(0, 40, n), g.integers(0, 40, n), g.integers(0, 16, n)]).astype(np.int32)
         for n in (30, 15)]
    d2.add_template_pyramid("c", [dict(width=40, height=40, tl_x=2, tl_y=4, features=f[0]),
                                  dict(width=20, height=20, tl_x=1, tl_y=2, features=f[1])])
    path = str(tmp_path / "c.yaml")
    d2.writeClass("c", path)
    d3 = sbm.Detector16(128, [4, 8])
    d3.readClass(path)
    for lvl in range(2):
        assert np.array_equal(d3.getTemplates("c", 0)[lvl].features, f[lvl])


def _bank_equal(a, b):
    assert a.classIds() == b.classIds()
    for cid in a.classIds():
        assert a.numTemplates(cid) == b.numTemplates(cid)
        for t in range(a.numTemplates(cid)):
            for la, lb in zip(a.getTemplates(cid, t), b.getTemplates(cid, t)):
                assert (la.width, la.height, la.tl_x, la.tl_y, la.pyramid_level) == \
                       (lb.width, lb.height, lb.tl_x, lb.tl_y, lb.pyramid_level)
                assert np.array_equal(la.features, lb.features)
                assert np.array_equal(la.theta.view(np.uint32), lb.theta.view(np.uint32))


@pytest.mark.parametrize("suffix", [".bank", ".bank.gz"])
def test_packed_bank_file_round_trip(tmp_path, suffix):
    det = sbm.Detector(63, [4, 8], 25.0, 55.0)
    det.readClass(golden(FIXTURE_BANKS[1][0]))
    g = np.random.default_rng(1)
    f = [np.column_stack([g.integers(-5, 40, n), g.integers(0, 40, n), g.integers(0, 8, n)]).astype(np.int32)
         for n in (40, 20)]
    th = [g.uniform(0, 360, len(x)).astype(np.float32) for x in f]
    det.add_template_pyramid("angles", [dict(width=40, height=40, tl_x=2, tl_y=4, features=f[0], theta=th[0]),
                                        dict(width=20, height=20, tl_x=1, tl_y=2, features=f[1], theta=th[1])])
    path = tmp_path / ("all" + suffix)
    det.save_bank(path)
    d2 = sbm.Detector([4, 8])
    assert d2.load_bank(path) == 2
    _bank_equal(det, d2)
    assert d2.getModalities() == dict(weak_threshold=25.0, num_features=63, strong_threshold=55.0)
    # the YAML written from the loaded bank is the reference's file, byte for byte
    det.writeClass("test", tmp_path / "a.yaml")
    d2.writeClass("test", tmp_path / "b.yaml")
    assert (tmp_path / "a.yaml").read_bytes() == (tmp_path / "b.yaml").read_bytes()


def test_packed_bank_file_errors(tmp_path):
    det = sbm.Detector(63, [4, 8])
    det.readClass(golden(FIXTURE_BANKS[2][0]))
    path = tmp_path / "b.bank"
    det.save_bank(path)
    with pytest.raises(sbm.SbmError) as e:  # another pyramid
        sbm.Detector([4, 8, 16]).load_bank(path)
    assert e.value.code == _lib.SBM_EINVAL
    with pytest.raises(sbm.SbmError) as e:  # 16 orientations
        sbm.Detector16([4, 8]).load_bank(path)
    assert e.value.code == _lib.SBM_EINVAL
    d2 = sbm.Detector([4, 8])
    d2.load_bank(path)
    with pytest.raises(sbm.SbmError) as e:  # the class is present
        d2.load_bank(path)
    assert e.value.code == _lib.SBM_EDUPCLASS and d2.numTemplates() == det.numTemplates()
    data = path.read_bytes()
    for bad in (data[:-3], b"NOTABANK" + data[8:], data + b"\0"):
        (tmp_path / "bad.bank").write_bytes(bad)
        d3 = sbm.Detector([4, 8])
        with pytest.raises(sbm.SbmError) as e:
            d3.load_bank(tmp_path / "bad.bank")
        assert e.value.code == _lib.SBM_EIO and d3.numClasses() == 0  # nothing half-loaded
    with pytest.raises(sbm.SbmError) as e:
        d3.load_bank(tmp_path / "missing.bank")
    assert e.value.code == _lib.SBM_EIO


def test_packed_bank_file_16_orientations(tmp_path):
    g = np.random.default_rng(2)
    det = sbm.Detector16(64, [4, 8])
    for k in range(3):
        f = [np.column_stack([g.integers(0, 60, n), g.integers(0, 60, n), g.integers(0, 16, n)]).astype(np.int32)
             for n in (64, 32)]
        det.add_template_pyramid("c%d" % (k % 2), [dict(width=60, height=60, tl_x=0, tl_y=0, features=f[0]),
                                                   dict(width=30, height=30, tl_x=0, tl_y=0, features=f[1])])
    det.save_bank(tmp_path / "b16.bank")
    d2 = sbm.Detector16(64, [4, 8])
    assert d2.load_bank(tmp_path / "b16.bank") == 2
    _bank_equal(det, d2)
    with pytest.raises(sbm.SbmError) as e:
        sbm.Detector(64, [4, 8]).load_bank(tmp_path / "b16.bank")
    assert e.value.code == _lib.SBM_EINVAL


def test_packed_bank_loads_faster_than_yaml(tmp_path):
    from shape_based_matching_amd import synth
    _, _, poly, _ = synth.scene_plain(0)
    det = sbm.Detector(128, [4, 8])
    for k in range(1500):
        det.add_template_pyramid("big", synth.bank_template(poly, k * 0.24, 0.9 + 0.1 * (k % 3), 60, 512))
    det.writeClass("big", tmp_path / "big.yaml")
    det.save_bank(tmp_path / "big.bank")
    t0 = time.perf_counter()
    sbm.Detector(128, [4, 8]).readClass(tmp_path / "big.yaml")
    t1 = time.perf_counter()
    d2 = sbm.Detector(128, [4, 8])
    d2.load_bank(tmp_path / "big.bank")
    t2 = time.perf_counter()
    print(f"1500 templates x 768 features: YAML {t1 - t0:.3f} s, packed {t2 - t1:.3f} s; "
          f"{(tmp_path / 'big.yaml').stat().st_size >> 20} MiB vs {(tmp_path / 'big.bank').stat().st_size >> 20} MiB")
    assert t2 - t1 < (t1 - t0) / 3
    _bank_equal(det, d2)


def test_no_cpu_fallback_without_gpu():
    n = ctypes.c_int(0)
    try:
        hip = ctypes.CDLL("libamdhip64.so")
        hip.hipGetDeviceCount(ctypes.byref(n))
    except OSError:
        pass
    if n.value > 0:
        pytest.skip("a GPU is visible")
    det = sbm.Detector(63, [4, 8])
    det.readClass(golden(FIXTURE_BANKS[2][0]))
    with pytest.raises(sbm.SbmError) as e:
        det.match(np.zeros((64, 64), np.uint8), 90)
    assert e.value.code == _lib.SBM_ENODEV
    with pytest.raises(sbm.SbmError) as e:
        sbm.Server(det, slots=2)
    assert e.value.code == _lib.SBM_ENODEV
    with pytest.raises(sbm.SbmError) as e:
        det.upload_bank()
    assert e.value.code == _lib.SBM_ENODEV


@pytest.mark.parametrize("seed", range(5))
def test_nms_matches_oracle(seed):
    rng = np.random.default_rng(seed)
    n = 60
    boxes = np.column_stack([rng.integers(0, 200, n), rng.integers(0, 200, n), rng.integers(5, 60, n),
                             rng.integers(5, 60, n)]).astype(np.int32)
    scores = rng.choice(np.array([0.0, 50.0, 80.5, 90.0, 95.25, 100.0], np.float32), n)
    for thr, eta, k in [(0.5, 1.0, 0), (0.3, 0.9, 0), (0.5, 1.0, 10)]:
        assert sbm.nms_boxes(boxes, scores, 0.0, thr, eta, k) == oracle.nms_boxes(boxes, scores, 0.0, thr, eta, k)


def test_match_ordering_semantics():
    a = sbm.Match(1, 2, 90.0, "c", 3)
    b = sbm.Match(1, 2, 90.0, "c", 4)
    c = sbm.Match(5, 5, 95.0, "c", 9)
    assert c < a < b and a == b and not (a == c)


@pytest.mark.parametrize("name", ["case1_test_info.yaml.gz", "case2_test_info.yaml.gz", "case0_circle_info.yaml.gz"])
def test_infos_yaml_round_trip_is_byte_exact(name, tmp_path):
    """load_infos then save_infos reproduces the reference's own info files byte for byte."""
    src = gzip.open(golden(name)).read()
    p = tmp_path / "in.yaml"
    p.write_bytes(src)
    infos = sbm.shapeInfo_producer.load_infos(str(p))
    assert len(infos) in (361, 89)
    out = tmp_path / "out.yaml"
    sbm.shapeInfo_producer.save_infos(infos, str(out))
    assert out.read_bytes() == src


@pytest.mark.parametrize("ranges", [((0, 360), (), 1.0, 0.5), ((), (0.1, 1.0), 15.0, 0.01), ((), (), 15.0, 0.5),
                                    ((-30, 30), (0.8, 1.2), 7.5, 0.1), ((45,), (0.5, 2.0), 15.0, 0.25)])
def test_produce_infos_equals_oracle(ranges):
    ar, sr, astep, sstep = ranges
    sp = sbm.shapeInfo_producer(np.zeros((8, 8), np.uint8))
    sp.angle_range, sp.scale_range, sp.angle_step, sp.scale_step = list(ar), list(sr), astep, sstep
    sp.produce_infos()
    ref = oracle.produce_infos(ar, sr, astep, sstep)
    assert [(i.angle, i.scale) for i in sp.infos] == [(float(np.float32(a)), float(np.float32(s))) for a, s in ref]
    assert len(sp.infos) >= 1


def test_shape_transform_needs_a_gpu():
    n = ctypes.c_int(0)
    try:
        ctypes.CDLL("libamdhip64.so").hipGetDeviceCount(ctypes.byref(n))
    except OSError:
        pass
    if n.value > 0:
        pytest.skip("a GPU is visible")
    with pytest.raises(sbm.SbmError) as e:
        sbm.shapeInfo_producer.transform(np.zeros((16, 16), np.uint8), 30.0, 1.0)
    assert e.value.code == _lib.SBM_ENODEV
