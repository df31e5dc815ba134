"""ctypes wrapper around ``liboracle.so`` (the C restatement in ``sbm_oracle.c``).

Test infrastructure only: the checker for the HIP path, and the CPU baseline of bench.py.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# SBM_ORACLE_LIBRARY selects another build of the oracle (the sanitized one of the CPU sanitizer run)
_LIB_PATH = os.environ.get("SBM_ORACLE_LIBRARY") or os.path.join(_HERE, "liboracle.so")
_lib = None

EINVAL = -1000000000


def build(force=False):
    """Compile liboracle.so with the Makefile next to this file."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.check_call(["make", "-s", "-B" if force else "-s", "-C", _HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = ctypes.CDLL(_LIB_PATH)
        _declare(_lib)
    return _lib


_u8p = ctypes.POINTER(ctypes.c_uint8)
_u16p = ctypes.POINTER(ctypes.c_uint16)
_i32p = ctypes.POINTER(ctypes.c_int32)
_f32p = ctypes.POINTER(ctypes.c_float)
_f64p = ctypes.POINTER(ctypes.c_double)
_c_int, _c_float = ctypes.c_int, ctypes.c_float


def _declare(L):
    sig = {
        "orc_gaussian_blur7": (None, [_u8p, _c_int, _c_int, _u8p]),
        "orc_sobel3": (None, [_u8p, _c_int, _c_int, _i32p, _i32p]),
        "orc_fast_atan2_deg": (_c_float, [_c_float, _c_float]),
        "orc_angle_bin": (_c_int, [_c_float]),
        "orc_bin_table": (None, [_c_int, _c_int, _u8p]),
        "orc_quantized_orientations": (None, [_u8p, _c_int, _c_int, _c_float, _u8p, _f32p, _f32p]),
        "orc_gradient_color": (None, [_u8p, _c_int, _c_int, _c_float, _u8p, _f32p, _f32p]),
        "orc_pyr_down": (None, [_u8p, _c_int, _c_int, _c_int, _u8p]),
        "orc_mask_down": (None, [_u8p, _c_int, _c_int, _u8p]),
        "orc_apply_mask": (None, [_u8p, _u8p, _c_int]),
        "orc_spread": (None, [_u8p, _c_int, _c_int, _c_int, _u8p]),
        "orc_response_maps": (None, [_u8p, _c_int, _c_int, _u8p]),
        "orc_similarity_lut": (ctypes.c_uint8, [_c_int, _c_int]),
        "orc_linearize": (None, [_u8p, _c_int, _c_int, _c_int, _u8p]),
        "orc_lm_guard_bytes": (_c_int, [_c_int, _c_int, _c_int]),
        "orc_build_lm_pyramid": (_c_int, [_u8p, _c_int, _c_int, _c_int, _u8p, _c_float, _c_int, _i32p,
                                          ctypes.POINTER(_u8p)]),
        "orc_similarity": (_c_int, [_u8p, _c_int, _c_int, _c_int, _i32p, _c_int, _c_int, _c_int, _u16p]),
        "orc_similarity_local": (_c_int, [_u8p, _c_int, _c_int, _c_int, _i32p, _c_int, _c_int, _c_int, _u16p]),
        "orc_match": (_c_int, [_u8p, _c_int, _c_int, _c_int, _u8p, _c_float, _c_int, _i32p, _c_float,
                               _c_int, _i32p, _i32p, _i32p, _i32p, _c_int, _i32p, _f32p, _c_int, _f64p]),
        "orc_match_lm": (_c_int, [ctypes.POINTER(_u8p), _c_int, _c_int, _c_int, _i32p, _c_float, _c_int, _i32p,
                                  _i32p, _i32p, _i32p, _c_int, _i32p, _f32p, _c_int]),
        "orc_score": (_c_float, [_c_int, _c_int]),
        "orc_sort_unique": (_c_int, [_c_int, _i32p, _f32p]),
        "orc_nms_boxes": (_c_int, [_c_int, _i32p, _f32p, _c_float, _c_float, _c_float, _c_int, _i32p]),
        "orc_extract_template": (_c_int, [_f32p, _u8p, _f32p, _u8p, _c_int, _c_int, _c_float, _c_int, _c_int,
                                          _i32p, _f32p]),
        "orc_extract_candidates": (_c_int, [_f32p, _u8p, _f32p, _u8p, _c_int, _c_int, _c_float, _c_int, _i32p,
                                            _f32p, _f32p]),
        "orc_add_template": (_c_int, [_u8p, _c_int, _c_int, _c_int, _u8p, _c_float, _c_float, _c_int, _c_int,
                                      _c_int, _i32p, _i32p, _f32p]),
        "orc_rotate_template": (None, [_c_int, _i32p, _i32p, _f32p, _c_float, _c_float, _c_float, _i32p, _i32p,
                                       _f32p]),
        "orc_rotation_matrix2d": (None, [_c_float, _c_float, ctypes.c_double, ctypes.c_double, _f64p]),
        "orc_invert_affine": (None, [_f64p, _f64p]),
        "orc_warp_affine": (None, [_u8p, _c_int, _c_int, _c_int, _f64p, _u8p]),
        "orc_shape_transform": (None, [_u8p, _c_int, _c_int, _c_int, _c_float, _c_float, _u8p]),
        "orc_produce_infos": (_c_int, [_c_int, _f32p, _c_int, _f32p, _c_float, _c_float, _c_float, _c_int, _f32p,
                                       _f32p]),
        # vectorised, OpenMP baseline (sbm_cpu_fast.c, chosen per CPU by sbm_cpu_fast_dispatch.c)
        "orc_fast_match": (_c_int, [_u8p, _c_int, _c_int, _c_int, _u8p, _c_float, _c_int, _i32p, _c_float,
                                    _c_int, _i32p, _i32p, _i32p, _i32p, _c_int, _i32p, _f32p, _c_int, _f64p]),
        "orc_fast_build_lm_pyramid": (_c_int, [_u8p, _c_int, _c_int, _c_int, _u8p, _c_float, _c_int, _i32p,
                                               ctypes.POINTER(_u8p)]),
        "orc_fast_isa": (ctypes.c_char_p, []),
        "orc_fast_bin_table": (None, [_c_int, _c_int, _u8p]),
        # 16-orientation variant (sbm_oracle16.c)
        "orc16_angle_bin": (_c_int, [_c_float]),
        "orc16_bin_table": (None, [_c_int, _c_int, _u8p]),
        "orc16_gradient_color": (None, [_u8p, _c_int, _c_int, _c_float, _u16p, _f32p, _f32p]),
        "orc16_apply_mask": (None, [_u16p, _u8p, _c_int]),
        "orc16_spread": (None, [_u16p, _c_int, _c_int, _c_int, _u16p]),
        "orc16_similarity_lut": (ctypes.c_uint8, [_c_int, _c_int]),
        "orc16_generated_lut": (ctypes.c_uint8, [_c_int, _c_int]),
        "orc16_response_maps": (None, [_u16p, _c_int, _c_int, _u8p]),
        "orc16_build_lm_pyramid": (_c_int, [_u8p, _c_int, _c_int, _c_int, _u8p, _c_float, _c_int, _i32p,
                                            ctypes.POINTER(_u8p)]),
        "orc16_similarity": (_c_int, [_u8p, _c_int, _c_int, _c_int, _i32p, _c_int, _c_int, _c_int, _u16p]),
        "orc16_similarity_local": (_c_int, [_u8p, _c_int, _c_int, _c_int, _i32p, _c_int, _c_int, _c_int, _u16p]),
        "orc16_match_lm": (_c_int, [ctypes.POINTER(_u8p), _c_int, _c_int, _c_int, _i32p, _c_float, _c_int, _i32p,
                                    _i32p, _i32p, _i32p, _c_int, _i32p, _f32p, _c_int]),
        "orc16_match": (_c_int, [_u8p, _c_int, _c_int, _c_int, _u8p, _c_float, _c_int, _i32p, _c_float, _c_int,
                                 _i32p, _i32p, _i32p, _i32p, _c_int, _i32p, _f32p, _c_int]),
        "orc16_extract_template": (_c_int, [_f32p, _u16p, _u8p, _c_int, _c_int, _c_float, _c_int, _c_int, _i32p]),
        "orc16_add_template": (_c_int, [_u8p, _c_int, _c_int, _c_int, _u8p, _c_float, _c_float, _c_int, _c_int,
                                        _c_int, _i32p, _i32p]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args


def _p(a, t):
    if a is None:
        return None
    return a.ctypes.data_as(t)


def _u8(a):
    return np.ascontiguousarray(a, dtype=np.uint8)


# ---------------------------------------------------------------------------------------------
# preprocessing
# ---------------------------------------------------------------------------------------------
def gaussian_blur7(img):
    img = _u8(img)
    out = np.empty_like(img)
    lib().orc_gaussian_blur7(_p(img, _u8p), img.shape[0], img.shape[1], _p(out, _u8p))
    return out


def sobel3(blurred):
    b = _u8(blurred)
    dx = np.empty(b.shape, np.int32)
    dy = np.empty(b.shape, np.int32)
    lib().orc_sobel3(_p(b, _u8p), b.shape[0], b.shape[1], _p(dx, _i32p), _p(dy, _i32p))
    return dx, dy


def fast_atan2_deg(y, x):
    return lib().orc_fast_atan2_deg(float(y), float(x))


def angle_bin(a):
    return lib().orc_angle_bin(float(a))


def bin_table(lo=-1020, hi=1020):
    n = hi - lo + 1
    out = np.empty((n, n), np.uint8)
    lib().orc_bin_table(lo, hi, _p(out, _u8p))
    return out


def quantized_orientations(img, weak=30.0):
    """Returns (quantized uint8, magnitude float32, angle float32) for a 1- or 3-channel image."""
    img = _u8(img)
    rows, cols = img.shape[:2]
    q = np.empty((rows, cols), np.uint8)
    mag = np.empty((rows, cols), np.float32)
    ag = np.empty((rows, cols), np.float32)
    if img.ndim == 2:
        lib().orc_quantized_orientations(_p(img, _u8p), rows, cols, weak, _p(q, _u8p), _p(mag, _f32p), _p(ag, _f32p))
    else:
        lib().orc_gradient_color(_p(img, _u8p), rows, cols, weak, _p(q, _u8p), _p(mag, _f32p), _p(ag, _f32p))
    return q, mag, ag


def pyr_down(img):
    img = _u8(img)
    ch = 1 if img.ndim == 2 else img.shape[2]
    rows, cols = img.shape[:2]
    shape = (rows // 2, cols // 2) if img.ndim == 2 else (rows // 2, cols // 2, ch)
    out = np.empty(shape, np.uint8)
    lib().orc_pyr_down(_p(img, _u8p), rows, cols, ch, _p(out, _u8p))
    return out


def mask_down(mask):
    m = _u8(mask)
    out = np.empty((m.shape[0] // 2, m.shape[1] // 2), np.uint8)
    lib().orc_mask_down(_p(m, _u8p), m.shape[0], m.shape[1], _p(out, _u8p))
    return out


def spread(q, T):
    q = _u8(q)
    out = np.empty_like(q)
    lib().orc_spread(_p(q, _u8p), q.shape[0], q.shape[1], T, _p(out, _u8p))
    return out


def response_maps(s):
    s = _u8(s)
    out = np.empty((8,) + s.shape, np.uint8)
    lib().orc_response_maps(_p(s, _u8p), s.shape[0], s.shape[1], _p(out, _u8p))
    return out


def similarity_lut():
    return np.array([lib().orc_similarity_lut(i // 32, i % 32) for i in range(256)], np.uint8)


def linearize(resp, T):
    r = _u8(resp)
    rows, cols = r.shape
    out = np.empty((T * T, (rows // T) * (cols // T)), np.uint8)
    lib().orc_linearize(_p(r, _u8p), rows, cols, T, _p(out, _u8p))
    return out


def level_shapes(rows, cols, num_levels):
    shapes = [(rows, cols)]
    for _ in range(1, num_levels):
        rows, cols = rows // 2, cols // 2
        shapes.append((rows, cols))
    return shapes


def lm_pyramid(img, T, weak=30.0, mask=None):
    """Linear memories per level, each a flat uint8 array [8][T*T][H*W] followed by a zero guard."""
    img = _u8(img)
    rows, cols = img.shape[:2]
    ch = 1 if img.ndim == 2 else img.shape[2]
    L = len(T)
    bufs = []
    for (r, c), t in zip(level_shapes(rows, cols, L), T):
        bufs.append(np.zeros(8 * r * c + lib().orc_lm_guard_bytes(r, c, t), np.uint8))
    arr = (_u8p * L)(*[_p(b, _u8p) for b in bufs])
    Ta = np.asarray(T, np.int32)
    m = None if mask is None else _u8(mask)
    rc = lib().orc_build_lm_pyramid(_p(img, _u8p), rows, cols, ch, _p(m, _u8p), weak, L, _p(Ta, _i32p), arr)
    if rc != 0:
        raise ValueError("image size not divisible at some pyramid level")
    return bufs


def similarity(lm, rows, cols, T, feats, width, height):
    f = np.ascontiguousarray(feats, np.int32).reshape(-1, 3)
    W, H = cols // T, rows // T
    out = np.empty((H, W), np.uint16)
    P = lib().orc_similarity(_p(lm, _u8p), rows, cols, T, _p(f, _i32p), len(f), width, height, _p(out, _u16p))
    if P == EINVAL:
        raise ValueError("feature size too large")
    return out


def similarity_local(lm, rows, cols, T, feats, cx, cy):
    f = np.ascontiguousarray(feats, np.int32).reshape(-1, 3)
    out = np.empty((16, 16), np.uint16)
    rc = lib().orc_similarity_local(_p(lm, _u8p), rows, cols, T, _p(f, _i32p), len(f), cx, cy, _p(out, _u16p))
    if rc == EINVAL:
        raise ValueError("feature size too large")
    return out


def score(raw, n):
    return lib().orc_score(int(raw), int(n))


# ---------------------------------------------------------------------------------------------
# matching
# ---------------------------------------------------------------------------------------------
def flatten_bank(class_templates, num_levels, class_ids=None):
    """class_templates: {class_id: [pyramid, ...]}, pyramid = [level dict(width,height,tl_x,tl_y,features)].

    Returns (tinfo, feats, tclass, tid, class_names) in the oracle's flat layout, following the
    reference's class order: std::map (sorted names) when class_ids is empty, else class_ids order
    (line2Dup.cpp:1095-1111).  Class ranks are positions in sorted-name order."""
    names = sorted(class_templates)
    rank = {n: i for i, n in enumerate(names)}
    order = names if not class_ids else [c for c in class_ids if c in class_templates]
    tinfo, feats, tclass, tid = [], [], [], []
    off = 0
    for cname in order:
        for t_id, pyr in enumerate(class_templates[cname]):
            if len(pyr) != num_levels:
                raise ValueError("template pyramid depth differs from the detector's")
            for lv in pyr:
                f = np.asarray(lv["features"], np.int32).reshape(-1, 3)
                tinfo.append([lv["width"], lv["height"], lv["tl_x"], lv["tl_y"], len(f), off])
                feats.append(f)
                off += len(f)
            tclass.append(rank[cname])
            tid.append(t_id)
    feats = np.concatenate(feats) if feats else np.zeros((0, 3), np.int32)
    return (np.asarray(tinfo, np.int32).reshape(-1, num_levels, 6), np.ascontiguousarray(feats, np.int32),
            np.asarray(tclass, np.int32), np.asarray(tid, np.int32), names)


def match(img, T, threshold, class_templates, class_ids=None, mask=None, weak=30.0, nthreads=1,
          max_out=1 << 20, return_times=False):
    """Detector::match on the CPU: list of (x, y, similarity, class_id, template_id) in canonical order."""
    img = _u8(img)
    rows, cols = img.shape[:2]
    ch = 1 if img.ndim == 2 else img.shape[2]
    L = len(T)
    tinfo, feats, tclass, tid, names = flatten_bank(class_templates, L, class_ids)
    Ta = np.asarray(T, np.int32)
    out_i = np.empty((max_out, 4), np.int32)
    out_s = np.empty(max_out, np.float32)
    times = np.zeros(2, np.float64)
    m = None if mask is None else _u8(mask)
    n = lib().orc_match(_p(img, _u8p), rows, cols, ch, _p(m, _u8p), weak, L, _p(Ta, _i32p), threshold,
                        len(tid), _p(tinfo, _i32p), _p(feats, _i32p), _p(tclass, _i32p), _p(tid, _i32p),
                        max_out, _p(out_i, _i32p), _p(out_s, _f32p), nthreads, _p(times, _f64p))
    if n == EINVAL:
        raise ValueError("invalid input (size not divisible at a pyramid level, or >= 8192 features)")
    if n < 0:
        return match(img, T, threshold, class_templates, class_ids, mask, weak, nthreads, -n, return_times)
    res = [(int(out_i[i, 0]), int(out_i[i, 1]), float(out_s[i]), names[out_i[i, 2]], int(out_i[i, 3]))
           for i in range(n)]
    if return_times:
        return res, (float(times[0]), float(times[1]))
    return res


def match_lm(lms, rows, cols, T, threshold, class_templates, class_ids=None, max_out=1 << 20, nthreads=1):
    """matchClass + sort/unique on given linear memories (lms[l]: flat [8][T*T][H*W] + guard)."""
    L = len(T)
    tinfo, feats, tclass, tid, names = flatten_bank(class_templates, L, class_ids)
    keep = [np.ascontiguousarray(b, np.uint8) for b in lms]
    arr = (_u8p * L)(*[_p(b, _u8p) for b in keep])
    Ta = np.asarray(T, np.int32)
    out_i = np.empty((max_out, 4), np.int32)
    out_s = np.empty(max_out, np.float32)
    n = lib().orc_match_lm(arr, rows, cols, L, _p(Ta, _i32p), threshold, len(tid), _p(tinfo, _i32p), _p(feats, _i32p),
                           _p(tclass, _i32p), _p(tid, _i32p), max_out, _p(out_i, _i32p), _p(out_s, _f32p), nthreads)
    if n == EINVAL:
        raise ValueError("invalid input")
    return [(int(out_i[i, 0]), int(out_i[i, 1]), float(out_s[i]), names[out_i[i, 2]], int(out_i[i, 3]))
            for i in range(n)]


def sort_unique(matches, class_names):
    """Canonical sort + std::unique over (x, y, similarity, class_id, template_id) tuples."""
    rank = {n: i for i, n in enumerate(sorted(class_names))}
    names = sorted(class_names)
    n = len(matches)
    i32 = np.zeros((max(n, 1), 4), np.int32)
    sim = np.zeros(max(n, 1), np.float32)
    for k, (x, y, s, c, t) in enumerate(matches):
        i32[k] = (x, y, rank[c], t)
        sim[k] = s
    m = lib().orc_sort_unique(n, _p(i32, _i32p), _p(sim, _f32p))
    return [(int(i32[k, 0]), int(i32[k, 1]), float(sim[k]), names[i32[k, 2]], int(i32[k, 3])) for k in range(m)]


def nms_boxes(boxes, scores, score_threshold, nms_threshold, eta=1.0, top_k=0):
    b = np.ascontiguousarray(boxes, np.int32).reshape(-1, 4)
    s = np.ascontiguousarray(scores, np.float32)
    idx = np.empty(max(len(b), 1), np.int32)
    k = lib().orc_nms_boxes(len(b), _p(b, _i32p), _p(s, _f32p), score_threshold, nms_threshold, eta, top_k,
                            _p(idx, _i32p))
    return [int(i) for i in idx[:k]]


# ---------------------------------------------------------------------------------------------
# training
# ---------------------------------------------------------------------------------------------
def extract_template(mag, angle, angle_ori, mask, strong, num_features):
    rows, cols = angle.shape
    mx = rows * cols
    out = np.empty((mx, 3), np.int32)
    th = np.empty(mx, np.float32)
    mag = np.ascontiguousarray(mag, np.float32)
    angle = _u8(angle)
    ao = np.ascontiguousarray(angle_ori, np.float32)
    m = None if mask is None else _u8(mask)
    n = lib().orc_extract_template(_p(mag, _f32p), _p(angle, _u8p), _p(ao, _f32p), _p(m, _u8p), rows, cols, strong,
                                   num_features, mx, _p(out, _i32p), _p(th, _f32p))
    if n < 0:
        return None
    return out[:n].copy(), th[:n].copy()


def extract_candidates(mag, angle, angle_ori, mask, strong=60.0):
    """The candidate stage of extractTemplate: (xyl int32 (n,3), theta, score), sorted by score."""
    rows, cols = angle.shape
    mx = rows * cols
    out = np.empty((mx, 3), np.int32)
    th = np.empty(mx, np.float32)
    sc = np.empty(mx, np.float32)
    mag = np.ascontiguousarray(mag, np.float32)
    angle = _u8(angle)
    ao = np.ascontiguousarray(angle_ori, np.float32)
    m = None if mask is None else _u8(mask)
    n = lib().orc_extract_candidates(_p(mag, _f32p), _p(angle, _u8p), _p(ao, _f32p), _p(m, _u8p), rows, cols, strong,
                                     mx, _p(out, _i32p), _p(th, _f32p), _p(sc, _f32p))
    return out[:n].copy(), th[:n].copy(), sc[:n].copy()


def add_template(img, num_features, num_levels=2, mask=None, weak=30.0, strong=60.0):
    """Detector::addTemplate: returns a pyramid (list of level dicts with features and theta) or None."""
    img = _u8(img)
    rows, cols = img.shape[:2]
    ch = 1 if img.ndim == 2 else img.shape[2]
    mx = rows * cols
    tinfo = np.zeros((num_levels, 6), np.int32)
    feats = np.zeros((mx, 3), np.int32)
    theta = np.zeros(mx, np.float32)
    m = None if mask is None else _u8(mask)
    n = lib().orc_add_template(_p(img, _u8p), rows, cols, ch, _p(m, _u8p), weak, strong, num_features, num_levels, mx,
                               _p(tinfo, _i32p), _p(feats, _i32p), _p(theta, _f32p))
    if n < 0:
        return None
    return _pyr_from_flat(tinfo, feats, theta)


def _pyr_from_flat(tinfo, feats, theta):
    pyr = []
    for lvl, row in enumerate(tinfo):
        w, h, tx, ty, n, off = (int(v) for v in row)
        pyr.append(dict(width=w, height=h, tl_x=tx, tl_y=ty, pyramid_level=lvl,
                        features=feats[off:off + n].copy(), theta=theta[off:off + n].copy()))
    return pyr


def rotate_template(pyr, theta, center):
    """Detector::addTemplate_rotate for one source pyramid (features need 'theta')."""
    L = len(pyr)
    tinfo = np.zeros((L, 6), np.int32)
    fl, tl = [], []
    off = 0
    for l, lv in enumerate(pyr):
        f = np.asarray(lv["features"], np.int32).reshape(-1, 3)
        tinfo[l] = (lv["width"], lv["height"], lv["tl_x"], lv["tl_y"], len(f), off)
        fl.append(f)
        tl.append(np.asarray(lv.get("theta", np.zeros(len(f))), np.float32))
        off += len(f)
    feats = np.ascontiguousarray(np.concatenate(fl), np.int32)
    th = np.ascontiguousarray(np.concatenate(tl), np.float32)
    dt = np.zeros_like(tinfo)
    df = np.zeros_like(feats)
    dth = np.zeros_like(th)
    lib().orc_rotate_template(L, _p(tinfo, _i32p), _p(feats, _i32p), _p(th, _f32p), float(theta), float(center[0]),
                              float(center[1]), _p(dt, _i32p), _p(df, _i32p), _p(dth, _f32p))
    return _pyr_from_flat(dt, df, dth)


# ---------------------------------------------------------------------------------------------
# shapeInfo_producer (line2Dup.h:209-332)
# ---------------------------------------------------------------------------------------------
def rotation_matrix2d(center, angle, scale):
    M = np.zeros(6, np.float64)
    lib().orc_rotation_matrix2d(float(center[0]), float(center[1]), float(angle), float(scale), _p(M, _f64p))
    return M.reshape(2, 3)


def invert_affine(M):
    M = np.ascontiguousarray(M, np.float64).reshape(6)
    out = np.zeros(6, np.float64)
    lib().orc_invert_affine(_p(M, _f64p), _p(out, _f64p))
    return out.reshape(2, 3)


def warp_affine(src, M):
    """cv::warpAffine(src, dst, M, src.size()), INTER_LINEAR, constant zero border."""
    src = _u8(src)
    rows, cols = src.shape[:2]
    ch = 1 if src.ndim == 2 else src.shape[2]
    M = np.ascontiguousarray(M, np.float64).reshape(6)
    dst = np.empty_like(src)
    lib().orc_warp_affine(_p(src, _u8p), rows, cols, ch, _p(M, _f64p), _p(dst, _u8p))
    return dst


def shape_transform(src, angle, scale):
    """shapeInfo_producer::transform (line2Dup.h:245-253)."""
    src = _u8(src)
    rows, cols = src.shape[:2]
    ch = 1 if src.ndim == 2 else src.shape[2]
    dst = np.empty_like(src)
    lib().orc_shape_transform(_p(src, _u8p), rows, cols, ch, float(angle), float(scale), _p(dst, _u8p))
    return dst


def mask_of(mask, angle, scale):
    """shapeInfo_producer::mask_of: transform(mask) > 0 as 0/255 (line2Dup.h:326-328)."""
    return np.where(shape_transform(mask, angle, scale) > 0, 255, 0).astype(np.uint8)


def produce_infos(angle_range=(), scale_range=(), angle_step=15.0, scale_step=0.5, eps=0.00001):
    """shapeInfo_producer::produce_infos (line2Dup.h:281-320): list of (angle, scale) float32 pairs."""
    ar = np.asarray(angle_range, np.float32)
    sr = np.asarray(scale_range, np.float32)
    n = lib().orc_produce_infos(len(ar), _p(ar, _f32p), len(sr), _p(sr, _f32p), float(angle_step), float(scale_step),
                                float(eps), 0, None, None)
    a = np.zeros(max(n, 1), np.float32)
    sc = np.zeros(max(n, 1), np.float32)
    lib().orc_produce_infos(len(ar), _p(ar, _f32p), len(sr), _p(sr, _f32p), float(angle_step), float(scale_step),
                            float(eps), n, _p(a, _f32p), _p(sc, _f32p))
    return list(zip(a[:n].tolist(), sc[:n].tolist()))


# ---------------------------------------------------------------------------------------------
# 16-orientation variant (test/ori_16bit_experiment/line2Dup_16bit_ori.cpp; sbm_oracle16.c)
# Defaults of that variant: weak 10, strong 55 (line2Dup_16bit_ori.cpp:500-512, 1051-1056).
# ---------------------------------------------------------------------------------------------
WEAK16, STRONG16 = 10.0, 55.0


def angle_bin16(a):
    return lib().orc16_angle_bin(float(a))


def bin_table16(lo=-1020, hi=1020):
    n = hi - lo + 1
    out = np.empty((n, n), np.uint8)
    lib().orc16_bin_table(lo, hi, _p(out, _u8p))
    return out


def quantized_orientations16(img, weak=WEAK16):
    """(quantized uint16 single bits, magnitude float32, angle float32) of a 3-channel BGR image."""
    img = _u8(img)
    if img.ndim != 3 or img.shape[2] != 3:
        raise ValueError("the 16-orientation variant reads three interleaved channels")
    rows, cols = img.shape[:2]
    q = np.empty((rows, cols), np.uint16)
    mag = np.empty((rows, cols), np.float32)
    ag = np.empty((rows, cols), np.float32)
    lib().orc16_gradient_color(_p(img, _u8p), rows, cols, weak, _p(q, _u16p), _p(mag, _f32p), _p(ag, _f32p))
    return q, mag, ag


def spread16(q, T):
    q = np.ascontiguousarray(q, np.uint16)
    out = np.empty_like(q)
    lib().orc16_spread(_p(q, _u16p), q.shape[0], q.shape[1], T, _p(out, _u16p))
    return out


def response_maps16(s):
    s = np.ascontiguousarray(s, np.uint16)
    out = np.empty((16,) + s.shape, np.uint8)
    lib().orc16_response_maps(_p(s, _u16p), s.shape[0], s.shape[1], _p(out, _u8p))
    return out


def similarity_lut16():
    return np.array([lib().orc16_similarity_lut(i // 64, i % 64) for i in range(1024)], np.uint8)


def generated_lut16():
    return np.array([lib().orc16_generated_lut(i // 64, i % 64) for i in range(1024)], np.uint8)


def lm_pyramid16(img, T, weak=WEAK16, mask=None):
    """Linear memories per level, each a flat uint8 array [16][T*T][H*W] followed by a zero guard."""
    img = _u8(img)
    rows, cols = img.shape[:2]
    L = len(T)
    bufs = []
    for (r, c), t in zip(level_shapes(rows, cols, L), T):
        bufs.append(np.zeros(16 * r * c + lib().orc_lm_guard_bytes(r, c, t), np.uint8))
    arr = (_u8p * L)(*[_p(b, _u8p) for b in bufs])
    Ta = np.asarray(T, np.int32)
    m = None if mask is None else _u8(mask)
    ch = 1 if img.ndim == 2 else img.shape[2]
    rc = lib().orc16_build_lm_pyramid(_p(img, _u8p), rows, cols, ch, _p(m, _u8p), weak, L, _p(Ta, _i32p), arr)
    if rc != 0:
        raise ValueError("not a 3-channel image, or size not divisible at some pyramid level")
    return bufs


def similarity16(lm, rows, cols, T, feats, width, height):
    f = np.ascontiguousarray(feats, np.int32).reshape(-1, 3)
    out = np.empty((rows // T, cols // T), np.uint16)
    P = lib().orc16_similarity(_p(lm, _u8p), rows, cols, T, _p(f, _i32p), len(f), width, height, _p(out, _u16p))
    if P == EINVAL:
        raise ValueError("feature size too large")
    return out


def similarity_local16(lm, rows, cols, T, feats, cx, cy):
    f = np.ascontiguousarray(feats, np.int32).reshape(-1, 3)
    out = np.empty((16, 16), np.uint16)
    rc = lib().orc16_similarity_local(_p(lm, _u8p), rows, cols, T, _p(f, _i32p), len(f), cx, cy, _p(out, _u16p))
    if rc == EINVAL:
        raise ValueError("feature size too large")
    return out


def match16(img, T, threshold, class_templates, class_ids=None, mask=None, weak=WEAK16, nthreads=1, max_out=1 << 20):
    """Detector::match of the 16-orientation variant: (x, y, similarity, class_id, template_id) tuples."""
    img = _u8(img)
    rows, cols = img.shape[:2]
    ch = 1 if img.ndim == 2 else img.shape[2]
    L = len(T)
    tinfo, feats, tclass, tid, names = flatten_bank(class_templates, L, class_ids)
    Ta = np.asarray(T, np.int32)
    out_i = np.empty((max_out, 4), np.int32)
    out_s = np.empty(max_out, np.float32)
    m = None if mask is None else _u8(mask)
    n = lib().orc16_match(_p(img, _u8p), rows, cols, ch, _p(m, _u8p), weak, L, _p(Ta, _i32p), threshold, len(tid),
                          _p(tinfo, _i32p), _p(feats, _i32p), _p(tclass, _i32p), _p(tid, _i32p), max_out,
                          _p(out_i, _i32p), _p(out_s, _f32p), nthreads)
    if n == EINVAL:
        raise ValueError("invalid input (not 3 channels, or size not divisible at a pyramid level)")
    if n < 0:
        return match16(img, T, threshold, class_templates, class_ids, mask, weak, nthreads, -n)
    return [(int(out_i[i, 0]), int(out_i[i, 1]), float(out_s[i]), names[out_i[i, 2]], int(out_i[i, 3]))
            for i in range(n)]


def match_lm16(lms, rows, cols, T, threshold, class_templates, class_ids=None, max_out=1 << 20, nthreads=1):
    """matchClass + sort/unique of the variant on given linear memories (lms[l]: [16][T*T][H*W] + guard)."""
    L = len(T)
    tinfo, feats, tclass, tid, names = flatten_bank(class_templates, L, class_ids)
    keep = [np.ascontiguousarray(b, np.uint8) for b in lms]
    arr = (_u8p * L)(*[_p(b, _u8p) for b in keep])
    Ta = np.asarray(T, np.int32)
    out_i = np.empty((max_out, 4), np.int32)
    out_s = np.empty(max_out, np.float32)
    n = lib().orc16_match_lm(arr, rows, cols, L, _p(Ta, _i32p), threshold, len(tid), _p(tinfo, _i32p),
                             _p(feats, _i32p), _p(tclass, _i32p), _p(tid, _i32p), max_out, _p(out_i, _i32p),
                             _p(out_s, _f32p), nthreads)
    if n == EINVAL:
        raise ValueError("invalid input")
    return [(int(out_i[i, 0]), int(out_i[i, 1]), float(out_s[i]), names[out_i[i, 2]], int(out_i[i, 3]))
            for i in range(n)]


def extract_template16(mag, angle, mask, strong, num_features):
    rows, cols = angle.shape
    mx = max(num_features, 1)
    out = np.empty((mx, 3), np.int32)
    mag = np.ascontiguousarray(mag, np.float32)
    angle = np.ascontiguousarray(angle, np.uint16)
    m = None if mask is None else _u8(mask)
    n = lib().orc16_extract_template(_p(mag, _f32p), _p(angle, _u16p), _p(m, _u8p), rows, cols, strong, num_features,
                                     mx, _p(out, _i32p))
    if n < 0:
        return None
    return out[:n].copy()


def add_template16(img, num_features, num_levels=2, mask=None, weak=WEAK16, strong=STRONG16):
    """Detector::addTemplate of the variant: a pyramid (list of level dicts) or None when refused."""
    img = _u8(img)
    rows, cols = img.shape[:2]
    ch = 1 if img.ndim == 2 else img.shape[2]
    mx = rows * cols
    tinfo = np.zeros((num_levels, 6), np.int32)
    feats = np.zeros((mx, 3), np.int32)
    m = None if mask is None else _u8(mask)
    n = lib().orc16_add_template(_p(img, _u8p), rows, cols, ch, _p(m, _u8p), weak, strong, num_features, num_levels,
                                 mx, _p(tinfo, _i32p), _p(feats, _i32p))
    if n < 0:
        return None
    return _pyr_from_flat(tinfo, feats, np.zeros(mx, np.float32))


# ---------------------------------------------------------------------------------------------
# the vectorised, OpenMP-parallel restatement (sbm_cpu_fast.c): the timed CPU baseline
# ---------------------------------------------------------------------------------------------
class FlatBank:
    """A bank in the oracle's flat layout, made once (flattening a large bank takes longer than matching)."""

    def __init__(self, class_templates, num_levels, class_ids=None):
        self.tinfo, self.feats, self.tclass, self.tid, self.names = flatten_bank(class_templates, num_levels,
                                                                                 class_ids)
        self.L = num_levels


def fast_isa():
    """Instruction set the fast baseline runs with on this CPU ("AVX-512", "AVX2" or "x86-64")."""
    return lib().orc_fast_isa().decode()


def match_fast(img, T, threshold, bank, mask=None, weak=30.0, nthreads=1, max_out=1 << 20):
    """Detector::match on the CPU, vectorised, OpenMP over templates (preprocessing on one thread, as the
    reference).  bank: FlatBank.  Returns (matches as match() gives them, (response-map seconds, match
    seconds))."""
    img = _u8(img)
    rows, cols = img.shape[:2]
    ch = 1 if img.ndim == 2 else img.shape[2]
    L = len(T)
    if bank.L != L:
        raise ValueError("bank pyramid depth differs from T")
    Ta = np.asarray(T, np.int32)
    out_i = np.empty((max_out, 4), np.int32)
    out_s = np.empty(max_out, np.float32)
    times = np.zeros(2, np.float64)
    m = None if mask is None else _u8(mask)
    n = lib().orc_fast_match(_p(img, _u8p), rows, cols, ch, _p(m, _u8p), weak, L, _p(Ta, _i32p), threshold,
                             len(bank.tid), _p(bank.tinfo, _i32p), _p(bank.feats, _i32p), _p(bank.tclass, _i32p),
                             _p(bank.tid, _i32p), max_out, _p(out_i, _i32p), _p(out_s, _f32p), nthreads,
                             _p(times, _f64p))
    if n == EINVAL:
        raise ValueError("invalid input (size not divisible at a pyramid level, or >= 8192 features)")
    if n < 0:
        return match_fast(img, T, threshold, bank, mask, weak, nthreads, -n)
    res = [(int(out_i[i, 0]), int(out_i[i, 1]), float(out_s[i]), bank.names[out_i[i, 2]], int(out_i[i, 3]))
           for i in range(n)]
    return res, (float(times[0]), float(times[1]))


def lm_pyramid_fast(img, T, weak=30.0, mask=None):
    """The fast baseline's linear memories (same layout as lm_pyramid)."""
    img = _u8(img)
    rows, cols = img.shape[:2]
    ch = 1 if img.ndim == 2 else img.shape[2]
    L = len(T)
    bufs, r, c = [], rows, cols
    for lvl in range(L):
        if lvl > 0:
            r, c = r // 2, c // 2
        bufs.append(np.zeros(8 * r * c + lib().orc_lm_guard_bytes(r, c, T[lvl]), np.uint8))
    arr = (_u8p * L)(*[_p(b, _u8p) for b in bufs])
    Ta = np.asarray(T, np.int32)
    m = None if mask is None else _u8(mask)
    rc = lib().orc_fast_build_lm_pyramid(_p(img, _u8p), rows, cols, ch, _p(m, _u8p), weak, L, _p(Ta, _i32p), arr)
    if rc != 0:
        raise ValueError("invalid input")
    return bufs


def fast_bin_table(lo=-1020, hi=1020):
    """The fast baseline's exhaustive (dx, dy) -> 8-bin table (layout of bin_table)."""
    n = hi - lo + 1
    out = np.empty((n, n), np.uint8)
    lib().orc_fast_bin_table(lo, hi, _p(out, _u8p))
    return out
