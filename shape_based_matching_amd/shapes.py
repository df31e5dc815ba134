"""shape_based_matching::shapeInfo_producer (line2Dup.h:209-332) over the C ABI.

``transform`` warps on the GPU (cv::getRotationMatrix2D about (cols/2, rows/2) + cv::warpAffine,
INTER_LINEAR, zero border, bit-exact with OpenCV's fixed-point warp); ``produce_infos`` and the
YAML ``save_infos`` / ``load_infos`` run in the library on the host.
"""
import ctypes

import numpy as np

from ._lib import check, lib

__all__ = ["Info", "shapeInfo_producer", "ShapeInfoProducer"]


class Info:
    """shapeInfo_producer::Info: one (angle, scale) pair, float32 as in the reference."""

    __slots__ = ("angle", "scale")

    def __init__(self, angle, scale):
        self.angle = float(np.float32(angle))
        self.scale = float(np.float32(scale))

    def __eq__(self, other):
        return isinstance(other, Info) and (self.angle, self.scale) == (other.angle, other.scale)

    __hash__ = None

    def __repr__(self):
        return f"Info(angle={self.angle!r}, scale={self.scale!r})"


def _image(a):
    a = np.ascontiguousarray(a, dtype=np.uint8)
    if a.ndim == 2:
        return a, a.shape[0], a.shape[1], 1
    if a.ndim == 3 and a.shape[2] in (1, 3):
        return a, a.shape[0], a.shape[1], a.shape[2]
    raise ValueError("images must be HxW (gray) or HxWx3 (BGR) uint8")


class shapeInfo_producer:  # noqa: N801 -- the reference's class name
    def __init__(self, src, mask=None, detector=None):
        self.src = np.ascontiguousarray(src, dtype=np.uint8)
        if mask is None:
            self.mask = np.full(self.src.shape[:2], 255, np.uint8)  # "make sure we have masks"
        else:
            self.mask = np.ascontiguousarray(mask, dtype=np.uint8)
        self.angle_range = []
        self.scale_range = []
        self.angle_step = 15.0
        self.scale_step = 0.5
        self.eps = 0.00001
        self.infos = []
        self._det = detector

    @staticmethod
    def transform(src, angle, scale, detector=None, binary_mask=False):
        a, rows, cols, ch = _image(src)
        out = np.empty_like(a)
        h = detector._h if detector is not None else None
        check(lib().sbm_shape_transform(h, a.ctypes.data, rows, cols, cols * ch, ch, float(angle), float(scale),
                                        1 if binary_mask else 0, out.ctypes.data))
        return out

    @staticmethod
    def save_infos(infos, path="infos.yaml"):
        a = np.array([i.angle for i in infos], np.float32)
        s = np.array([i.scale for i in infos], np.float32)
        check(lib().sbm_save_infos(str(path).encode(), a.ctypes.data if len(a) else None,
                                   s.ctypes.data if len(s) else None, len(a)))

    @staticmethod
    def load_infos(path="info.yaml"):
        cnt = ctypes.c_int()
        check(lib().sbm_load_infos(str(path).encode(), None, None, 0, ctypes.byref(cnt)))
        n = cnt.value
        a = np.zeros(max(n, 1), np.float32)
        s = np.zeros(max(n, 1), np.float32)
        check(lib().sbm_load_infos(str(path).encode(), a.ctypes.data, s.ctypes.data, n, ctypes.byref(cnt)))
        return [Info(x, y) for x, y in zip(a[:n], s[:n])]

    def produce_infos(self):
        ar = np.array(self.angle_range, np.float32)
        sr = np.array(self.scale_range, np.float32)
        cnt = ctypes.c_int()
        args = (ar.ctypes.data if len(ar) else None, len(ar), sr.ctypes.data if len(sr) else None, len(sr),
                float(self.angle_step), float(self.scale_step), float(self.eps))
        check(lib().sbm_produce_infos(*args, None, None, 0, ctypes.byref(cnt)))
        n = cnt.value
        a = np.zeros(max(n, 1), np.float32)
        s = np.zeros(max(n, 1), np.float32)
        check(lib().sbm_produce_infos(*args, a.ctypes.data, s.ctypes.data, n, ctypes.byref(cnt)))
        self.infos = [Info(x, y) for x, y in zip(a[:n], s[:n])]

    def src_of(self, info):
        return self.transform(self.src, info.angle, info.scale, self._det)

    def mask_of(self, info):
        return self.transform(self.mask, info.angle, info.scale, self._det, binary_mask=True)


ShapeInfoProducer = shapeInfo_producer
