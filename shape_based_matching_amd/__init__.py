"""MI355X-native line2Dup matching path (meiqua/shape_based_matching's Detector::match).

The product is ``libsbm_hip.so`` (C ABI in ``include/sbm_hip.h``); this package is its Python
binding, with names mirroring the reference's ``line2Dup`` API.
"""
from ._lib import LIB_PATH, SbmError, lib  # noqa: F401
from .detector import (MODE_BATCHED, MODE_PLAIN, Detector, Detector16, Feature, Match, Server,  # noqa: F401
                       Template, nms_boxes)
from .shapes import Info, ShapeInfoProducer, shapeInfo_producer  # noqa: F401

__version__ = "0.1.0"
