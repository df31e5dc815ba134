"""Python mirror of the reference's ``line2Dup`` API (line2Dup.h:13-205, test.cpp:114-119).

Every call goes through the C ABI of ``libsbm_hip.so``; the matching path runs on the GPU.
Method names follow the reference (``addTemplate``, ``addTemplate_rotate``, ``match``,
``readClasses``...), so tests read like code written against the reference.
"""
import ctypes
import functools

import numpy as np

from ._lib import MODE_BATCHED, MODE_PLAIN, SbmError, check, lib

__all__ = ["Feature", "Template", "Match", "Detector", "Detector16", "Server", "nms_boxes", "MODE_PLAIN", "MODE_BATCHED",
           "SbmError"]


class Feature:
    """Feature{x, y, label, theta} (line2Dup.h:13-26)."""
    __slots__ = ("x", "y", "label", "theta")

    def __init__(self, x=0, y=0, label=0, theta=0.0):
        self.x, self.y, self.label, self.theta = int(x), int(y), int(label), float(theta)

    def __repr__(self):
        return f"Feature({self.x}, {self.y}, {self.label})"


class Template:
    """Template{width, height, tl_x, tl_y, pyramid_level, features} (line2Dup.h:28-39).

    ``features`` is an (n, 3) int32 array of (x, y, label); ``theta`` an (n,) float32 array."""

    def __init__(self, width=0, height=0, tl_x=0, tl_y=0, pyramid_level=0, features=None, theta=None):
        self.width, self.height, self.tl_x, self.tl_y = int(width), int(height), int(tl_x), int(tl_y)
        self.pyramid_level = int(pyramid_level)
        self.features = np.zeros((0, 3), np.int32) if features is None else np.asarray(features, np.int32).reshape(-1, 3)
        self.theta = (np.zeros(len(self.features), np.float32) if theta is None
                      else np.asarray(theta, np.float32).reshape(-1))

    def as_dict(self):
        return dict(width=self.width, height=self.height, tl_x=self.tl_x, tl_y=self.tl_y,
                    pyramid_level=self.pyramid_level, features=self.features, theta=self.theta)


@functools.total_ordering
class Match:
    """Match{x, y, similarity, class_id, template_id} with the reference's ordering and equality
    (line2Dup.h:108-141)."""
    __slots__ = ("x", "y", "similarity", "class_id", "template_id")

    def __init__(self, x, y, similarity, class_id, template_id):
        self.x, self.y, self.similarity, self.class_id, self.template_id = x, y, similarity, class_id, template_id

    def __lt__(self, other):
        if self.similarity != other.similarity:
            return self.similarity > other.similarity
        return self.template_id < other.template_id

    def __eq__(self, other):
        return (self.x == other.x and self.y == other.y and self.similarity == other.similarity
                and self.class_id == other.class_id)

    __hash__ = None

    def astuple(self):
        return (self.x, self.y, self.similarity, self.class_id, self.template_id)

    def __repr__(self):
        return f"Match(x={self.x}, y={self.y}, similarity={self.similarity}, class_id={self.class_id!r}, " \
               f"template_id={self.template_id})"


def _img_args(img):
    a = np.ascontiguousarray(img, dtype=np.uint8)
    if a.ndim == 2:
        rows, cols, ch = a.shape[0], a.shape[1], 1
    elif a.ndim == 3 and a.shape[2] in (1, 3):
        rows, cols, ch = a.shape
    else:
        raise ValueError("images must be HxW (gray) or HxWx3 (BGR) uint8")
    return a, rows, cols, cols * ch, ch


def _ids(class_ids):
    ids = [c.encode() for c in (class_ids or [])]
    arr = (ctypes.c_char_p * max(len(ids), 1))(*ids)
    return arr, len(ids)


class DeviceImage:
    """An image resident in HBM (sbm_image)."""

    def __init__(self, handle, shape):
        self._h = handle
        self.shape = shape

    def close(self):
        if getattr(self, "_h", None):
            lib().sbm_image_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Detector:
    """line2Dup::Detector (line2Dup.h:143-205).

    ``Detector()``: T = {4, 8}, 63 features, thresholds 30/60.  ``Detector(T=[...])`` or
    ``Detector(num_features, T, weak_thresh=30, strong_thresh=60)`` as the reference's three
    constructors (line2Dup.cpp:1027-1047)."""

    def __init__(self, num_features=None, T=None, weak_thresh=30.0, strong_thresh=60.0):
        if isinstance(num_features, (list, tuple)) and T is None:
            num_features, T = None, num_features
        T = [4, 8] if T is None else [int(t) for t in T]
        arr = (ctypes.c_int * len(T))(*T)
        h = ctypes.c_void_p()
        check(lib().sbm_detector_create(int(num_features or 0), arr, len(T), float(weak_thresh), float(strong_thresh),
                                        ctypes.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            lib().sbm_detector_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- getters (line2Dup.h:163-175) -----------------------------------------------------------
    def pyramidLevels(self):
        v = ctypes.c_int()
        check(lib().sbm_pyramid_levels(self._h, ctypes.byref(v)))
        return v.value

    def getT(self, pyramid_level):
        v = ctypes.c_int()
        check(lib().sbm_get_T(self._h, int(pyramid_level), ctypes.byref(v)))
        return v.value

    def getModalities(self):
        w, n, s = ctypes.c_float(), ctypes.c_int(), ctypes.c_float()
        check(lib().sbm_get_modality(self._h, ctypes.byref(w), ctypes.byref(n), ctypes.byref(s)))
        return dict(weak_threshold=w.value, num_features=n.value, strong_threshold=s.value)

    def orientationBins(self):
        """8, or 16 for the 16-orientation variant (Detector16)."""
        v = ctypes.c_int()
        check(lib().sbm_orientation_bins(self._h, ctypes.byref(v)))
        return v.value

    def numTemplates(self, class_id=None):
        v = ctypes.c_int()
        check(lib().sbm_num_templates(self._h, None if class_id is None else class_id.encode(), ctypes.byref(v)))
        return v.value

    def numClasses(self):
        v = ctypes.c_int()
        check(lib().sbm_num_classes(self._h, ctypes.byref(v)))
        return v.value

    def classIds(self):
        out = []
        buf = ctypes.create_string_buffer(4096)
        for i in range(self.numClasses()):
            check(lib().sbm_class_id(self._h, i, buf, len(buf)))
            out.append(buf.value.decode())
        return out

    def _class_names(self):
        """classIds(), cached: classes are never removed, so an unchanged count means unchanged names."""
        n = self.numClasses()
        cached = getattr(self, "_names_cache", None)
        if cached is None or len(cached) != n:
            cached = self.classIds()
            self._names_cache = cached
        return cached

    def getTemplates(self, class_id, template_id):
        out = []
        for lvl in range(self.pyramidLevels()):
            v = [ctypes.c_int() for _ in range(6)]
            check(lib().sbm_get_template(self._h, class_id.encode(), int(template_id), lvl, *[ctypes.byref(x) for x in v]))
            n = v[5].value
            f = np.zeros((n, 3), np.int32)
            th = np.zeros(n, np.float32)
            check(lib().sbm_get_template_features(self._h, class_id.encode(), int(template_id), lvl,
                                                  f.ctypes.data, th.ctypes.data, n))
            out.append(Template(v[0].value, v[1].value, v[2].value, v[3].value, v[4].value, f, th))
        return out

    # ---- bank building ---------------------------------------------------------------------------
    def addTemplate(self, sources, class_id, object_mask=None, num_features=0):
        """Detector::addTemplate (line2Dup.cpp:1270-1304); returns the id or -1."""
        a, rows, cols, stride, ch = _img_args(sources)
        m = None if object_mask is None else np.ascontiguousarray(object_mask, np.uint8)
        if m is not None and m.shape != (rows, cols):
            raise ValueError("mask size differs from the image")
        tid = ctypes.c_int()
        check(lib().sbm_add_template(self._h, a.ctypes.data, rows, cols, stride, ch,
                                     None if m is None else m.ctypes.data, class_id.encode(), int(num_features),
                                     ctypes.byref(tid)))
        return tid.value

    def addTemplate_rotate(self, class_id, zero_id, theta, center):
        """Detector::addTemplate_rotate (line2Dup.cpp:1320-1362)."""
        tid = ctypes.c_int()
        check(lib().sbm_add_template_rotate(self._h, class_id.encode(), int(zero_id), float(theta), float(center[0]),
                                            float(center[1]), ctypes.byref(tid)))
        return tid.value

    def add_template_pyramid(self, class_id, pyramid):
        """Add a template pyramid given as Template objects or dicts (what readClass builds)."""
        L = len(pyramid)
        tinfo = np.zeros((L, 6), np.int32)
        feats, thetas, off = [], [], 0
        for lvl, t in enumerate(pyramid):
            d = t.as_dict() if isinstance(t, Template) else t
            f = np.asarray(d["features"], np.int32).reshape(-1, 3)
            tinfo[lvl] = (d["width"], d["height"], d["tl_x"], d["tl_y"], len(f), off)
            feats.append(f)
            th = d.get("theta")
            thetas.append(np.zeros(len(f), np.float32) if th is None else np.asarray(th, np.float32))
            off += len(f)
        f = np.ascontiguousarray(np.concatenate(feats) if feats else np.zeros((0, 3)), np.int32)
        th = np.ascontiguousarray(np.concatenate(thetas) if thetas else np.zeros(0), np.float32)
        tid = ctypes.c_int()
        check(lib().sbm_add_template_features(self._h, class_id.encode(), L, tinfo.ctypes.data,
                                              f.ctypes.data if len(f) else None, th.ctypes.data if len(th) else None,
                                              ctypes.byref(tid)))
        return tid.value

    # ---- matching --------------------------------------------------------------------------------
    def match_arrays(self, sources, threshold, class_ids=None, masks=None, mode=MODE_BATCHED):
        """Detector::match returning (xyct int32 (n,4), similarity float32 (n,), class names)."""
        a, rows, cols, stride, ch = _img_args(sources)
        m = None if masks is None else np.ascontiguousarray(masks, np.uint8)
        if m is not None and m.shape != (rows, cols):
            raise SbmError(-1, "mask size differs from the image (line2Dup.cpp:1057)")
        ids, n_ids = _ids(class_ids)
        out = ctypes.c_void_p()
        check(lib().sbm_match(self._h, a.ctypes.data, rows, cols, stride, ch, float(threshold), ids, n_ids,
                              None if m is None else m.ctypes.data, int(mode), ctypes.byref(out)))
        return self._collect(out)

    def upload_image(self, img, mask=None):
        """Copy an image (and mask) once into HBM; match it with ``match_image``."""
        a, rows, cols, stride, ch = _img_args(img)
        m = None if mask is None else np.ascontiguousarray(mask, np.uint8)
        h = ctypes.c_void_p()
        check(lib().sbm_image_create(self._h, a.ctypes.data, rows, cols, stride, ch,
                                     None if m is None else m.ctypes.data, ctypes.byref(h)))
        return DeviceImage(h, (rows, cols, ch))

    def match_image_arrays(self, dimg, threshold, class_ids=None, mode=MODE_BATCHED):
        ids, n_ids = _ids(class_ids)
        out = ctypes.c_void_p()
        check(lib().sbm_match_image(self._h, dimg._h, float(threshold), ids, n_ids, int(mode), ctypes.byref(out)))
        return self._collect(out)

    def match_image(self, dimg, threshold, class_ids=None, mode=MODE_BATCHED):
        xyct, sim, names = self.match_image_arrays(dimg, threshold, class_ids, mode)
        return [Match(int(r[0]), int(r[1]), float(s), names[r[2]], int(r[3])) for r, s in zip(xyct, sim)]

    def match_device_noalloc_arrays(self, d_img, rows, cols, stride, channels, threshold, class_ids=None,
                                    d_mask=None, cap=4096):
        """sbm_match_device_noalloc on a device pointer (an int address on the detector's GPU): the batched
        match with every buffer in the library's fixed device array.  Returns (xyct, similarity, names)."""
        ids, n_ids = _ids(class_ids)
        xyct = np.zeros((max(cap, 1), 4), np.int32)
        sim = np.zeros(max(cap, 1), np.float32)
        cnt = ctypes.c_int()
        check(lib().sbm_match_device_noalloc(self._h, ctypes.c_void_p(d_img), rows, cols, stride, channels,
                                             float(threshold), ids, n_ids,
                                             None if d_mask is None else ctypes.c_void_p(d_mask),
                                             xyct.ctypes.data, sim.ctypes.data, cap, ctypes.byref(cnt)))
        n = min(cnt.value, cap)
        return xyct[:n], sim[:n], self._class_names(), cnt.value

    def match_images_arrays(self, dimgs, threshold, class_ids=None, slots=4):
        """Batched match of several resident images through per-size captured graphs (several images per
        launch).  Returns one (xyct, similarity, names) triple per image."""
        n = len(dimgs)
        arr = (ctypes.c_void_p * max(n, 1))(*[d._h.value if isinstance(d._h, ctypes.c_void_p) else d._h for d in dimgs])
        ids, n_ids = _ids(class_ids)
        counts = np.zeros(max(n, 1), np.int32)
        cap = max(getattr(self, "_batch_cap", 0), 256 * max(n, 1))
        while True:
            xyct = np.empty((cap, 4), np.int32)
            sim = np.empty(cap, np.float32)
            check(lib().sbm_match_batch_arrays(self._h, arr, n, float(threshold), ids, n_ids, int(slots),
                                               xyct.ctypes.data, sim.ctypes.data, cap, counts.ctypes.data))
            total = int(counts[:n].sum())
            if total <= cap:
                break
            cap = total  # a larger list than the arrays held: match again into larger arrays
        self._batch_cap = cap
        names = self._class_names()
        res, at = [], 0
        for c in counts[:n]:
            res.append((xyct[at:at + c], sim[at:at + c], names))
            at += int(c)
        return res

    def match_host_images_arrays(self, imgs, threshold, class_ids=None, masks=None, slots=4):
        """match_images_arrays for images in host memory (staged through pinned memory, one DMA per launch)."""
        n = len(imgs)
        arrs = [_img_args(im) for im in imgs]
        ms = None
        if masks is not None:
            ms = [None if m is None else np.ascontiguousarray(m, np.uint8) for m in masks]
        ptrs = (ctypes.c_void_p * max(n, 1))(*[a[0].ctypes.data for a in arrs])
        mptrs = None if ms is None else (ctypes.c_void_p * max(n, 1))(*[None if m is None else m.ctypes.data
                                                                         for m in ms])
        rows = np.array([a[1] for a in arrs] or [0], np.int32)
        cols = np.array([a[2] for a in arrs] or [0], np.int32)
        strides = np.array([a[3] for a in arrs] or [0], np.int32)
        chans = np.array([a[4] for a in arrs] or [1], np.int32)
        ids, n_ids = _ids(class_ids)
        counts = np.zeros(max(n, 1), np.int32)
        cap = max(getattr(self, "_batch_cap", 0), 256 * max(n, 1))
        while True:
            xyct = np.empty((cap, 4), np.int32)
            sim = np.empty(cap, np.float32)
            check(lib().sbm_match_batch_host_arrays(self._h, ptrs, rows.ctypes.data, cols.ctypes.data,
                                                    strides.ctypes.data, chans.ctypes.data, mptrs, n, float(threshold),
                                                    ids, n_ids, int(slots), xyct.ctypes.data, sim.ctypes.data, cap,
                                                    counts.ctypes.data))
            total = int(counts[:n].sum())
            if total <= cap:
                break
            cap = total
        self._batch_cap = cap
        names = self._class_names()
        res, at = [], 0
        for c in counts[:n]:
            res.append((xyct[at:at + c], sim[at:at + c], names))
            at += int(c)
        return res

    def set_images_per_launch(self, n):
        """Images of one size bucket per launch in match_images (1..16; 0: half a bucket)."""
        check(lib().sbm_set_images_per_launch(self._h, int(n)))

    def match_images(self, dimgs, threshold, class_ids=None, slots=4):
        out = []
        for xyct, sim, names in self.match_images_arrays(dimgs, threshold, class_ids, slots):
            out.append([Match(int(r[0]), int(r[1]), float(s), names[r[2]], int(r[3])) for r, s in zip(xyct, sim)])
        return out

    def add_templates_transformed(self, src, class_id, infos, mask=None, num_features=None):
        """addTemplate(shapes.src_of(info), class_id, shapes.mask_of(info), num_features[i]) for every
        info of a shapeInfo_producer over ``src`` / ``mask``, in order (warps and gradients on the GPU).
        ``infos``: Info objects or (angle, scale) pairs.  Returns the template ids (-1 where refused)."""
        a, rows, cols, stride, ch = _img_args(src)
        m = None if mask is None else np.ascontiguousarray(mask, np.uint8)
        if m is not None and m.shape != (rows, cols):
            raise SbmError(-1, "mask size differs from the image")
        pairs = [(i.angle, i.scale) if hasattr(i, "angle") else (i[0], i[1]) for i in infos]
        n = len(pairs)
        ang = np.array([p[0] for p in pairs], np.float32)
        sc = np.array([p[1] for p in pairs], np.float32)
        nf = None
        if num_features is not None:
            nf = np.broadcast_to(np.asarray(num_features, np.int32), (n,)).copy()
        ids = np.full(max(n, 1), -1, np.int32)
        check(lib().sbm_add_templates_transformed(self._h, a.ctypes.data, rows, cols, stride, ch,
                                                  None if m is None else m.ctypes.data, class_id.encode(),
                                                  ang.ctypes.data, sc.ctypes.data, None if nf is None else nf.ctypes.data,
                                                  n, ids.ctypes.data))
        return [int(v) for v in ids[:n]]

    def synchronize(self):
        check(lib().sbm_synchronize(self._h))

    def _collect(self, out):
        try:
            cnt = ctypes.c_int()
            check(lib().sbm_matches_count(out, ctypes.byref(cnt)))
            n = cnt.value
            xyct = np.zeros((max(n, 1), 4), np.int32)
            sim = np.zeros(max(n, 1), np.float32)
            got = ctypes.c_int()
            check(lib().sbm_matches_copy(out, xyct.ctypes.data, sim.ctypes.data, n, ctypes.byref(got)))
            names = self._class_names()
        finally:
            lib().sbm_matches_free(out)
        return xyct[:n], sim[:n], names

    def match(self, sources, threshold, class_ids=None, masks=None, mode=MODE_BATCHED):
        """Detector::match (line2Dup.cpp:1049-1121): list of Match in the canonical order."""
        xyct, sim, names = self.match_arrays(sources, threshold, class_ids, masks, mode)
        return [Match(int(r[0]), int(r[1]), float(s), names[r[2]], int(r[3])) for r, s in zip(xyct, sim)]

    def match_nms_arrays(self, source, threshold, class_ids=None, score_threshold=0.0, nms_threshold=0.5, top_k=0,
                         mask=None, with_all=False):
        """match followed by NMSBoxes over (x, y, templ[0].width, templ[0].height) as the demos do
        (test.cpp:428-451), the list and NMS built on the device.  ``source`` is an image array or a
        DeviceImage.  Returns (xyct, similarity, names) of the kept matches in NMS order, and with
        ``with_all`` also the full match list's triple."""
        ids, n_ids = _ids(class_ids)
        st = float(score_threshold)
        kept, full = ctypes.c_void_p(), ctypes.c_void_p()
        fp = ctypes.byref(full) if with_all else None
        if isinstance(source, DeviceImage):
            if mask is not None:
                raise SbmError(-1, "a DeviceImage carries its own mask")
            check(lib().sbm_match_image_nms(self._h, source._h, float(threshold), ids, n_ids, st, float(nms_threshold),
                                            int(top_k), ctypes.byref(kept), fp))
        else:
            a, rows, cols, stride, ch = _img_args(source)
            m = None if mask is None else np.ascontiguousarray(mask, np.uint8)
            if m is not None and m.shape != (rows, cols):
                raise SbmError(-1, "mask size differs from the image (line2Dup.cpp:1057)")
            check(lib().sbm_match_nms(self._h, a.ctypes.data, rows, cols, stride, ch,
                                      None if m is None else m.ctypes.data, float(threshold), ids, n_ids, st,
                                      float(nms_threshold), int(top_k), ctypes.byref(kept), fp))
        k = self._collect(kept)
        return (k, self._collect(full)) if with_all else k

    def match_nms(self, source, threshold, class_ids=None, score_threshold=0.0, nms_threshold=0.5, top_k=0,
                  mask=None):
        """Kept matches (list of Match, NMS order) of match + NMSBoxes; the defaults are the demos'
        (score_threshold 0, nms_threshold 0.5, test.cpp:450)."""
        xyct, sim, names = self.match_nms_arrays(source, threshold, class_ids, score_threshold, nms_threshold,
                                                 top_k, mask)
        return [Match(int(r[0]), int(r[1]), float(s), names[r[2]], int(r[3])) for r, s in zip(xyct, sim)]

    # ---- persistence (line2Dup.h:177-185) --------------------------------------------------------
    def readClasses(self, class_ids, format="templates_%s.yml.gz"):
        ids, n = _ids(class_ids)
        check(lib().sbm_read_classes(self._h, ids, n, format.encode()))

    def writeClasses(self, format="templates_%s.yml.gz"):
        check(lib().sbm_write_classes(self._h, format.encode()))

    def readClass(self, path, class_id_override=""):
        buf = ctypes.create_string_buffer(4096)
        check(lib().sbm_read_class_file(self._h, str(path).encode(), class_id_override.encode(), buf, len(buf)))
        return buf.value.decode()

    def writeClass(self, class_id, path):
        check(lib().sbm_write_class_file(self._h, class_id.encode(), str(path).encode()))

    def save_bank(self, path):
        """The whole bank and the detector's parameters as one packed binary file (sbm_save_bank)."""
        check(lib().sbm_save_bank(self._h, str(path).encode()))

    def upload_bank(self):
        """Packs the bank and copies it to HBM now rather than at the next match (sbm_upload_bank)."""
        check(lib().sbm_upload_bank(self._h))

    def load_bank(self, path):
        """Adds the classes of a packed binary bank file; returns how many (sbm_load_bank)."""
        n = ctypes.c_int()
        check(lib().sbm_load_bank(self._h, str(path).encode(), ctypes.byref(n)))
        self._names_cache = None
        return n.value

    def write(self, path):
        check(lib().sbm_write_detector_file(self._h, str(path).encode()))

    def read(self, path):
        check(lib().sbm_read_detector_file(self._h, str(path).encode()))

    # ---- tuning / instrumentation ----------------------------------------------------------------
    def set_tuning(self, positions_per_lane=16, waves_per_block=4, blocks_per_cu=8, gather=2, split=1):
        """Coarse tile shape (see sbm_set_tuning): gather 0 aligned loads realigned in registers, 1 unaligned
        vector loads, 2 aligned vectors + DPP, 3 nibble-packed (32 positions per lane); split > 1 lets that
        many wavefronts share one work item."""
        check(lib().sbm_set_tuning(self._h, positions_per_lane, waves_per_block, blocks_per_cu, gather, split))

    def get_tuning(self):
        """(positions_per_lane, waves_per_block, blocks_per_cu, gather, split)."""
        v = [ctypes.c_int() for _ in range(5)]
        check(lib().sbm_get_tuning(self._h, *[ctypes.byref(x) for x in v]))
        return tuple(x.value for x in v)

    def autotune(self, dimg, threshold, class_ids=None, iters=2, batch=None):
        """Pick the fastest launch shapes for a resident image: the coarse stage on the path the device takes,
        then the refinement.  batch (resident images of dimg's size and channels, unmasked): the multi-image
        launches of set_images_per_launch take dimg and these in turn instead of copies of dimg.  Returns (the
        gather kernel's shape, rows of every stage: dicts with "stage" in "gather", "list", "refine", the stage's
        parameters and its time in microseconds)."""
        ids, n = _ids(class_ids)
        v = [ctypes.c_int() for _ in range(5)]
        if batch:
            ims = [dimg] + list(batch)
            arr = (ctypes.c_void_p * len(ims))(*[d._h.value if isinstance(d._h, ctypes.c_void_p) else d._h
                                                  for d in ims])
            check(lib().sbm_autotune_images(self._h, arr, len(ims), float(threshold), ids, n, int(iters)))
            check(lib().sbm_get_tuning(self._h, *[ctypes.byref(x) for x in v]))
        else:
            cnt = ctypes.c_int()
            check(lib().sbm_autotune(self._h, dimg._h, float(threshold), ids, n, int(iters),
                                     *[ctypes.byref(x) for x in v], None, 0, ctypes.byref(cnt)))
        total = ctypes.c_int()
        check(lib().sbm_autotune_row(self._h, -1, None, None, None, ctypes.byref(total)))
        keys = {0: ("positions_per_lane", "waves_per_block", "blocks_per_cu", "gather", "split"),
                1: ("list_threads", "list_segment", "list_blocks_per_cu", "list_order", "list_split"),
                2: ("refine_waves", "refine_group", "refine_blocks_per_cu", "refine_layout")}
        rows = []
        for i in range(total.value):
            st, us = ctypes.c_int(), ctypes.c_float()
            p = np.zeros(5, np.int32)
            check(lib().sbm_autotune_row(self._h, i, ctypes.byref(st), p.ctypes.data, ctypes.byref(us), None))
            row = {"stage": ("gather", "list", "refine")[st.value], "us": float(us.value)}
            row.update({k: int(p[j]) for j, k in enumerate(keys[st.value])})
            rows.append(row)
        return tuple(x.value for x in v), rows

    def set_tuning_param(self, name, value):
        """One named launch-shape parameter (see sbm_set_tuning_param)."""
        check(lib().sbm_set_tuning_param(self._h, name.encode(), int(value)))

    def get_tuning_param(self, name):
        v = ctypes.c_int()
        check(lib().sbm_get_tuning_param(self._h, name.encode(), ctypes.byref(v)))
        return v.value

    def set_timing(self, enabled=True):
        """True / 1: per-kernel HIP events (eager launches); 2: only each graph replay's device span
        (read with last_device_ms); False / 0: off."""
        check(lib().sbm_set_timing(self._h, int(enabled) if enabled in (0, 1, 2) else 1))

    def set_coarse_mode(self, mode):
        """Batched coarse stage: 0 gather kernel, 1 (default) nonzero-response path on sparse coarse
        levels (at most 5% nonzero), 2 nonzero-response path whenever its list fits."""
        check(lib().sbm_set_coarse_mode(self._h, int(mode)))

    def get_coarse_mode(self):
        m = ctypes.c_int()
        check(lib().sbm_get_coarse_mode(self._h, ctypes.byref(m)))
        return m.value

    def last_coarse_path(self):
        """(ran the nonzero-response path, nonzero responses of the coarse level or -1)."""
        sp, nz = ctypes.c_int(), ctypes.c_int()
        check(lib().sbm_last_coarse_path(self._h, ctypes.byref(sp), ctypes.byref(nz)))
        return bool(sp.value), nz.value

    def last_coarse_entries(self):
        """List entries the last batched match's nonzero-response path read (0 when it did not run)."""
        n = ctypes.c_longlong()
        check(lib().sbm_last_coarse_entries(self._h, ctypes.byref(n)))
        return n.value

    def last_refine_features(self):
        """Features of the candidates the last match's finest-level refinement scored, summed."""
        v = ctypes.c_longlong()
        check(lib().sbm_last_refine_features(self._h, ctypes.byref(v)))
        return v.value

    def last_device_ms(self):
        ms = ctypes.c_float()
        check(lib().sbm_last_device_ms(self._h, ctypes.byref(ms)))
        return ms.value

    def last_timing(self):
        p, c, r = ctypes.c_float(), ctypes.c_float(), ctypes.c_float()
        cc, fc = ctypes.c_int(), ctypes.c_int()
        check(lib().sbm_last_timing(self._h, ctypes.byref(p), ctypes.byref(c), ctypes.byref(r), ctypes.byref(cc),
                                    ctypes.byref(fc)))
        return dict(preprocess_ms=p.value, coarse_ms=c.value, refine_ms=r.value, coarse_candidates=cc.value,
                    final_candidates=fc.value)

    def last_kernel_times(self):
        """[(kernel name, ms)] of the last match when timing is enabled."""
        n = ctypes.c_int()
        check(lib().sbm_last_kernel_time(self._h, -1, None, 0, None, ctypes.byref(n)))
        out = []
        buf = ctypes.create_string_buffer(64)
        for i in range(n.value):
            ms = ctypes.c_float()
            check(lib().sbm_last_kernel_time(self._h, i, buf, len(buf), ctypes.byref(ms), ctypes.byref(n)))
            out.append((buf.value.decode(), ms.value))
        return out

    def coarse_gather_bytes(self, rows, cols, class_ids=None):
        ids, n = _ids(class_ids)
        v = ctypes.c_longlong()
        check(lib().sbm_coarse_gather_bytes(self._h, rows, cols, ids, n, ctypes.byref(v)))
        return v.value

    def copy_probe(self, nbytes):
        """One plain copy kernel of nbytes (a known-size reference for traffic counters)."""
        check(lib().sbm_copy_probe(self._h, int(nbytes)))

    def device_bytes(self):
        v = ctypes.c_longlong()
        check(lib().sbm_device_bytes(self._h, ctypes.byref(v)))
        return v.value

    # ---- stage access (parity tests) -------------------------------------------------------------
    def prepare_image(self, img, mask=None):
        a, rows, cols, stride, ch = _img_args(img)
        m = None if mask is None else np.ascontiguousarray(mask, np.uint8)
        check(lib().sbm_prepare_image(self._h, a.ctypes.data, rows, cols, stride, ch,
                                      None if m is None else m.ctypes.data))
        self._prepared = (rows, cols, ch)

    def level_shape(self, level):
        rows, cols, _ = self._prepared
        for _ in range(level):
            rows, cols = rows // 2, cols // 2
        return rows, cols

    def read_quantized(self, level):
        r, c = self.level_shape(level)
        out = np.zeros((r, c), np.uint16 if self.orientationBins() == 16 else np.uint8)
        check(lib().sbm_read_level(self._h, level, 0, out.ctypes.data))
        return out

    def read_linear_memories(self, level):
        """[planes][T*T][H*W] bytes (planes = orientationBins()), without the guard."""
        r, c = self.level_shape(level)
        out = np.zeros(self.orientationBins() * r * c, np.uint8)
        check(lib().sbm_read_level(self._h, level, 1, out.ctypes.data))
        return out

    def read_nibbles(self, level):
        """The coarse level's nibble-packed linear memories (two positions per byte, guard included);
        present after prepare_image when the coarse gather is 3."""
        r, c = self.level_shape(level)
        T = self.getT(level)
        W, H = c // T, r // T
        guard = ((W * H + 17 * W + 4096) + 15) & ~15
        n = ((self.orientationBins() * r * c + guard) // 2 + 15) & ~15
        out = np.zeros(n, np.uint8)
        check(lib().sbm_read_level(self._h, level, 3, out.ctypes.data))
        return out

    def read_band(self, level):
        """A refined level's banded copy as the last match made it: with refine_layout 1 or 3
        [planes*T*T][ceil(H/16)][W+16][32] bytes, with refine_layout 4 (the nibble-banded copy) 16-byte slots,
        then the 1024-byte zeroed block."""
        r, c = self.level_shape(level)
        T = self.getT(level)
        W, H = c // T, r // T
        slot = 16 if self.get_tuning_param("refine_layout") == 4 else 32
        n = self.orientationBins() * T * T * ((H + 15) // 16) * (W + 16) * slot + 1024
        out = np.zeros(n, np.uint8)
        check(lib().sbm_read_level(self._h, level, 4, out.ctypes.data))
        return out

    def read_level_image(self, level):
        r, c = self.level_shape(level)
        ch = self._prepared[2]
        out = np.zeros((r, c, ch) if ch > 1 else (r, c), np.uint8)
        check(lib().sbm_read_level(self._h, level, 2, out.ctypes.data))
        return out

    def similarity_map(self, class_id, template_id):
        L = self.pyramidLevels()
        r, c = self.level_shape(L - 1)
        T = self.getT(L - 1)
        out = np.zeros((r // T, c // T), np.uint16)
        check(lib().sbm_similarity_map(self._h, class_id.encode(), int(template_id), out.ctypes.data))
        return out

    def similarity_local(self, class_id, template_id, level, cx, cy):
        out = np.zeros((16, 16), np.uint16)
        check(lib().sbm_similarity_local(self._h, class_id.encode(), int(template_id), int(level), int(cx), int(cy),
                                         out.ctypes.data))
        return out

    def gradient(self, img):
        a, rows, cols, stride, ch = _img_args(img)
        q = np.zeros((rows, cols), np.uint16 if self.orientationBins() == 16 else np.uint8)
        mag = np.zeros((rows, cols), np.float32)
        ang = np.zeros((rows, cols), np.float32)
        check(lib().sbm_gradient(self._h, a.ctypes.data, rows, cols, stride, ch, q.ctypes.data, mag.ctypes.data,
                                 ang.ctypes.data))
        return q, mag, ang

    def bin_table(self, lo=-1020, hi=1020):
        n = hi - lo + 1
        out = np.zeros((n, n), np.uint8)
        check(lib().sbm_bin_table(self._h, lo, hi, out.ctypes.data))
        return out

    def division_check(self):
        """Quotients of the gradient kernels' phase step that differ from IEEE division over their domain."""
        n = ctypes.c_int(-1)
        check(lib().sbm_division_check(self._h, ctypes.byref(n)))
        return n.value


class Detector16(Detector):
    """line2Dup::Detector of the 16-orientation variant (test/ori_16bit_experiment/line2Dup_16bit_ori.cpp).

    ``Detector16()``: T = {5, 8}, 63 features; ``Detector16(T)``; ``Detector16(num_features, T)`` -- the
    variant's three constructors (line2Dup_16bit_ori.cpp:1036-1056), with its fixed thresholds weak 10 and
    strong 55.  Sixteen orientations over 180 degrees, 3-channel (BGR) images only, templates refused
    below num_features candidates, no addTemplate_rotate; every other method is the Detector's."""

    def __init__(self, num_features=None, T=None):
        if isinstance(num_features, (list, tuple)) and T is None:
            num_features, T = None, num_features
        T = [5, 8] if T is None else [int(t) for t in T]
        arr = (ctypes.c_int * len(T))(*T)
        h = ctypes.c_void_p()
        check(lib().sbm_detector_create16(int(num_features or 0), arr, len(T), ctypes.byref(h)))
        self._h = h


def nms_boxes(boxes, scores, score_threshold, nms_threshold, eta=1.0, top_k=0):
    """cv_dnn::NMSBoxes (test.cpp:114-119): indices of the kept boxes, boxes as (x, y, w, h)."""
    b = np.ascontiguousarray(boxes, np.int32).reshape(-1, 4)
    s = np.ascontiguousarray(scores, np.float32).reshape(-1)
    idx = np.zeros(max(len(b), 1), np.int32)
    cnt = ctypes.c_int()
    check(lib().sbm_nms_boxes(b.ctypes.data if len(b) else None, s.ctypes.data if len(s) else None, len(b),
                              float(score_threshold), float(nms_threshold), float(eta), int(top_k),
                              idx.ctypes.data, ctypes.byref(cnt)))
    return [int(i) for i in idx[:cnt.value]]


class Server:
    """Continuous-batching request loop over a detector's resident bank (sbm_server).

    ``submit`` queues an image and returns a request id at once; ``wait`` returns its matches (the
    same list ``Detector.match`` gives for that image alone) and the latency from submission."""

    def __init__(self, detector, slots=4, hbm_budget_bytes=16 << 30):
        h = ctypes.c_void_p()
        check(lib().sbm_server_create(detector._h, int(slots), int(hbm_budget_bytes), ctypes.byref(h)))
        self._h = h
        self._det = detector  # keeps the detector alive

    def close(self):
        if getattr(self, "_h", None):
            lib().sbm_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, img, threshold, class_ids=None, mask=None):
        a, rows, cols, stride, ch = _img_args(img)
        m = None if mask is None else np.ascontiguousarray(mask, np.uint8)
        ids, n = _ids(class_ids)
        rid = ctypes.c_longlong()
        check(lib().sbm_server_submit(self._h, a.ctypes.data, rows, cols, stride, ch,
                                      None if m is None else m.ctypes.data, float(threshold), ids, n, ctypes.byref(rid)))
        return rid.value

    def wait_arrays(self, request_id, timeout_ms=-1):
        out = ctypes.c_void_p()
        lat = ctypes.c_double()
        check(lib().sbm_server_wait(self._h, int(request_id), int(timeout_ms), ctypes.byref(out), ctypes.byref(lat)))
        return self._det._collect(out), lat.value

    def wait(self, request_id, timeout_ms=-1):
        (xyct, sim, names), lat = self.wait_arrays(request_id, timeout_ms)
        return [Match(int(r[0]), int(r[1]), float(s), names[r[2]], int(r[3])) for r, s in zip(xyct, sim)], lat

    def stats(self):
        v = [ctypes.c_longlong() for _ in range(6)]
        check(lib().sbm_server_stats(self._h, *[ctypes.byref(x) for x in v]))
        keys = ("submitted", "admitted", "completed", "failed", "deferrals", "peak_bytes")
        out = dict(zip(keys, (x.value for x in v)))
        n = ctypes.c_longlong()
        check(lib().sbm_server_launches(self._h, ctypes.byref(n)))
        out["launches"] = n.value
        return out
