"""sbm_match_device_noalloc: the batched match of a device-resident image with every buffer in the
library's fixed device array returns exactly what sbm_match_device returns, on consecutive calls of
different sizes (the second smaller, then the first again), gray and colour, with and without a mask;
it refuses a bank that is not in HBM yet and an image larger than the array."""
import ctypes

import numpy as np
import pytest

import shape_based_matching_amd as sbm
from shape_based_matching_amd import _lib, synth

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


class DeviceBuffer:
    """Device memory for the test's inputs, through the HIP runtime the library itself links."""

    _hip = None

    def __init__(self, arr):
        if DeviceBuffer._hip is None:
            h = ctypes.CDLL("libamdhip64.so.7")
            h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
            h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
            h.hipFree.argtypes = [ctypes.c_void_p]
            h.hipDeviceSynchronize.argtypes = []
            DeviceBuffer._hip = h
        a = np.ascontiguousarray(arr, np.uint8)
        self.p = ctypes.c_void_p()
        assert DeviceBuffer._hip.hipMalloc(ctypes.byref(self.p), a.nbytes) == 0
        assert DeviceBuffer._hip.hipMemcpy(self.p, a.ctypes.data, a.nbytes, 1) == 0  # host to device

    def free(self):
        if self.p:
            DeviceBuffer._hip.hipDeviceSynchronize()
            DeviceBuffer._hip.hipFree(self.p)
            self.p = None


def match_device(det, buf, img, thr, mbuf=None):
    """The existing entry point, sbm_match_device, batched mode."""
    rows, cols = img.shape[:2]
    ch = 1 if img.ndim == 2 else img.shape[2]
    ids = (ctypes.c_char_p * 1)()
    out = ctypes.c_void_p()
    _lib.check(_lib.lib().sbm_match_device(det._h, buf.p, rows, cols, cols * ch, ch, float(thr), ids, 0,
                                           None if mbuf is None else mbuf.p, sbm.MODE_BATCHED, ctypes.byref(out)))
    xyct, sim, _ = det._collect(out)
    return xyct, sim


def noalloc(det, buf, img, thr, mbuf=None, cap=4096):
    rows, cols = img.shape[:2]
    ch = 1 if img.ndim == 2 else img.shape[2]
    xyct, sim, _, n = det.match_device_noalloc_arrays(buf.p.value, rows, cols, cols * ch, ch, thr,
                                                      d_mask=None if mbuf is None else mbuf.p.value, cap=cap)
    assert n == len(sim)
    return xyct, sim


def same(a, b):
    return np.array_equal(a[0], b[0]) and np.array_equal(a[1].view(np.uint32), b[1].view(np.uint32))


@pytest.fixture(scope="module")
def det():
    _, _, poly, _ = synth.scene_plain(0)
    d = sbm.Detector(256, [4, 8])
    for a in range(0, 360, 2):
        d.add_template_pyramid("poly", synth.bank_template(poly, float(a), 1.0, 60, 256))
    for a in range(0, 360, 6):
        d.add_template_pyramid("small", synth.bank_template(poly, float(a), 0.9, 60, 48))
    d.upload_bank()
    return d


def test_two_calls_second_smaller_equal_match_device(det):
    big = synth.scene_copies(1)[0]                      # 1024 x 1024
    small, _, _, _ = synth.scene_plain(0)               # 480 x 640
    rows, cols = small.shape
    mask = np.zeros((rows, cols), np.uint8)
    mask[32:-32, 48:-48] = 255
    bufs = [DeviceBuffer(big), DeviceBuffer(small), DeviceBuffer(mask)]
    try:
        ref_big = match_device(det, bufs[0], big, 80)
        ref_small = match_device(det, bufs[1], small, 70, bufs[2])
        assert len(ref_big[1]) > 0 and len(ref_small[1]) > 0
        got_big = noalloc(det, bufs[0], big, 80)
        got_small = noalloc(det, bufs[1], small, 70, bufs[2])   # the second call, smaller, masked
        say("1024x1024:", len(got_big[1]), "matches; 480x640 masked:", len(got_small[1]))
        assert same(got_big, ref_big)
        assert same(got_small, ref_small)
        assert same(noalloc(det, bufs[0], big, 80), ref_big)  # leftovers of the smaller call do not matter
        # a cap below the count copies the first matches in order and still reports them all
        xyct, sim, _, n = det.match_device_noalloc_arrays(bufs[0].p.value, 1024, 1024, 1024, 1, 80, cap=3)
        assert n == len(ref_big[1]) and np.array_equal(xyct, ref_big[0][:3])
    finally:
        for b in bufs:
            b.free()


def test_other_tunings_and_colour(det):
    img = synth.scene_copies(2)[0]
    bgr = np.ascontiguousarray(np.dstack([img, img[:, ::-1], 255 - img]))
    bufs = [DeviceBuffer(img), DeviceBuffer(bgr)]
    saved = det.get_tuning()
    try:
        for tuning in ((16, 4, 8, 2, 1), (32, 4, 4, 3, 4), (8, 2, 8, 0, 1)):
            det.set_tuning(*tuning)
            for buf, im in zip(bufs, (img, bgr)):
                ref = match_device(det, buf, im, 75)
                got = noalloc(det, buf, im, 75)
                say("tuning", tuning, "channels", 1 if im.ndim == 2 else 3, "matches", len(got[1]))
                assert same(got, ref)
    finally:
        det.set_tuning(*saved)
        for b in bufs:
            b.free()


def test_refusals(det):
    _, _, poly, _ = synth.scene_plain(0)
    fresh = sbm.Detector(256, [4, 8])
    fresh.add_template_pyramid("poly", synth.bank_template(poly, 0.0, 1.0, 60, 64))
    img = synth.scene_plain(3)[0]
    buf = DeviceBuffer(img)
    huge = DeviceBuffer(np.zeros((4096, 4096), np.uint8))
    try:
        with pytest.raises(sbm.SbmError) as e:  # bank not in HBM yet
            noalloc(fresh, buf, img, 80)
        assert e.value.code == _lib.SBM_EINVAL
        fresh.upload_bank()
        assert noalloc(fresh, buf, img, 80)[1].dtype == np.float32
        with pytest.raises(sbm.SbmError) as e:  # larger than the fixed array
            noalloc(det, huge, np.zeros((4096, 4096), np.uint8), 80)
        assert e.value.code == _lib.SBM_EBUDGET
        with pytest.raises(sbm.SbmError) as e:  # size not divisible at every level
            det.match_device_noalloc_arrays(buf.p.value, 470, 640, 640, 1, 80)
        assert e.value.code == _lib.SBM_ESIZE
        # the detector still matches normally afterwards
        assert same(noalloc(det, buf, img, 80), match_device(det, buf, img, 80))
    finally:
        buf.free()
        huge.free()
        fresh.close()
