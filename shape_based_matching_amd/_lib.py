"""ctypes binding of libsbm_hip.so (the C ABI in include/sbm_hip.h).

The library is built in-tree (``python -m shape_based_matching_amd.build``).  Loading fails loudly
when it is missing: there is no CPU fallback for the matching path.
"""
import ctypes
import os
import re

HERE = os.path.dirname(os.path.abspath(__file__))
# SBM_HIP_LIBRARY selects another build of the same library (the host-sanitized one of the CPU sanitizer run)
LIB_PATH = os.environ.get("SBM_HIP_LIBRARY") or os.path.join(HERE, "libsbm_hip.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "sbm_hip.h")

SBM_OK = 0
SBM_EINVAL = -1
SBM_ESIZE = -2
SBM_EFEATURES = -3
SBM_EDUPCLASS = -4
SBM_EHIP = -5
SBM_EBUDGET = -6
SBM_EIO = -7
SBM_ENOTFOUND = -8
SBM_ENODEV = -10
SBM_ETIMEOUT = -11
MODE_PLAIN = 0
MODE_BATCHED = 1

_c_int, _c_float = ctypes.c_int, ctypes.c_float
_vp = ctypes.c_void_p
_cp = ctypes.c_char_p
_ip = ctypes.POINTER(ctypes.c_int)
_fp = ctypes.POINTER(ctypes.c_float)
_llp = ctypes.POINTER(ctypes.c_longlong)
_pp = ctypes.POINTER(ctypes.c_void_p)
_cpp = ctypes.POINTER(ctypes.c_char_p)

_SIGS = {
    "sbm_last_error": (_cp, []),
    "sbm_version": (_cp, []),
    "sbm_detector_create": (_c_int, [_c_int, _ip, _c_int, _c_float, _c_float, _pp]),
    "sbm_detector_create16": (_c_int, [_c_int, _ip, _c_int, _pp]),
    "sbm_orientation_bins": (_c_int, [_vp, _ip]),
    "sbm_save_bank": (_c_int, [_vp, _cp]),
    "sbm_load_bank": (_c_int, [_vp, _cp, _ip]),
    "sbm_upload_bank": (_c_int, [_vp]),
    "sbm_detector_destroy": (_c_int, [_vp]),
    "sbm_pyramid_levels": (_c_int, [_vp, _ip]),
    "sbm_get_T": (_c_int, [_vp, _c_int, _ip]),
    "sbm_get_modality": (_c_int, [_vp, _fp, _ip, _fp]),
    "sbm_num_templates": (_c_int, [_vp, _cp, _ip]),
    "sbm_num_classes": (_c_int, [_vp, _ip]),
    "sbm_class_id": (_c_int, [_vp, _c_int, _cp, _c_int]),
    "sbm_get_template": (_c_int, [_vp, _cp, _c_int, _c_int, _ip, _ip, _ip, _ip, _ip, _ip]),
    "sbm_get_template_features": (_c_int, [_vp, _cp, _c_int, _c_int, _vp, _vp, _c_int]),
    "sbm_add_template": (_c_int, [_vp, _vp, _c_int, _c_int, _c_int, _c_int, _vp, _cp, _c_int, _ip]),
    "sbm_add_template_rotate": (_c_int, [_vp, _cp, _c_int, _c_float, _c_float, _c_float, _ip]),
    "sbm_add_template_features": (_c_int, [_vp, _cp, _c_int, _vp, _vp, _vp, _ip]),
    "sbm_match": (_c_int, [_vp, _vp, _c_int, _c_int, _c_int, _c_int, _c_float, _cpp, _c_int, _vp, _c_int, _pp]),
    "sbm_match_device": (_c_int, [_vp, _vp, _c_int, _c_int, _c_int, _c_int, _c_float, _cpp, _c_int, _vp, _c_int,
                                  _pp]),
    "sbm_match_device_noalloc": (_c_int, [_vp, _vp, _c_int, _c_int, _c_int, _c_int, _c_float, _cpp, _c_int, _vp,
                                          _vp, _vp, _c_int, _ip]),
    "sbm_image_create": (_c_int, [_vp, _vp, _c_int, _c_int, _c_int, _c_int, _vp, _pp]),
    "sbm_image_destroy": (_c_int, [_vp]),
    "sbm_match_image": (_c_int, [_vp, _vp, _c_float, _cpp, _c_int, _c_int, _pp]),
    "sbm_match_nms": (_c_int, [_vp, _vp, _c_int, _c_int, _c_int, _c_int, _vp, _c_float, _cpp, _c_int, _c_float,
                               _c_float, _c_int, _pp, _pp]),
    "sbm_match_image_nms": (_c_int, [_vp, _vp, _c_float, _cpp, _c_int, _c_float, _c_float, _c_int, _pp, _pp]),
    "sbm_match_batch": (_c_int, [_vp, _vp, _c_int, _c_float, _cpp, _c_int, _c_int, _vp]),
    "sbm_synchronize": (_c_int, [_vp]),
    "sbm_set_images_per_launch": (_c_int, [_vp, _c_int]),
    "sbm_match_batch_arrays": (_c_int, [_vp, _vp, _c_int, _c_float, _cpp, _c_int, _c_int, _vp, _vp, _c_int, _vp]),
    "sbm_match_batch_host_arrays": (_c_int, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _c_int, _c_float, _cpp, _c_int, _c_int,
                                             _vp, _vp, _c_int, _vp]),
    "sbm_matches_count": (_c_int, [_vp, _ip]),
    "sbm_matches_get": (_c_int, [_vp, _c_int, _ip, _ip, _fp, _ip, _ip]),
    "sbm_matches_copy": (_c_int, [_vp, _vp, _vp, _c_int, _ip]),
    "sbm_matches_class_id": (_c_int, [_vp, _c_int, _cp, _c_int]),
    "sbm_matches_free": (_c_int, [_vp]),
    "sbm_shape_transform": (_c_int, [_vp, _vp, _c_int, _c_int, _c_int, _c_int, _c_float, _c_float, _c_int, _vp]),
    "sbm_produce_infos": (_c_int, [_vp, _c_int, _vp, _c_int, _c_float, _c_float, _c_float, _vp, _vp, _c_int, _ip]),
    "sbm_save_infos": (_c_int, [_cp, _vp, _vp, _c_int]),
    "sbm_load_infos": (_c_int, [_cp, _vp, _vp, _c_int, _ip]),
    "sbm_add_templates_transformed": (_c_int, [_vp, _vp, _c_int, _c_int, _c_int, _c_int, _vp, _cp, _vp, _vp, _vp, _c_int,
                                               _vp]),
    "sbm_server_create": (_c_int, [_vp, _c_int, ctypes.c_longlong, _pp]),
    "sbm_server_destroy": (_c_int, [_vp]),
    "sbm_server_submit": (_c_int, [_vp, _vp, _c_int, _c_int, _c_int, _c_int, _vp, _c_float, _cpp, _c_int, _llp]),
    "sbm_server_wait": (_c_int, [_vp, ctypes.c_longlong, _c_int, _pp, ctypes.POINTER(ctypes.c_double)]),
    "sbm_server_stats": (_c_int, [_vp, _llp, _llp, _llp, _llp, _llp, _llp]),
    "sbm_server_launches": (_c_int, [_vp, _llp]),
    "sbm_read_classes": (_c_int, [_vp, _cpp, _c_int, _cp]),
    "sbm_write_classes": (_c_int, [_vp, _cp]),
    "sbm_read_class_file": (_c_int, [_vp, _cp, _cp, _cp, _c_int]),
    "sbm_write_class_file": (_c_int, [_vp, _cp, _cp]),
    "sbm_write_detector_file": (_c_int, [_vp, _cp]),
    "sbm_read_detector_file": (_c_int, [_vp, _cp]),
    "sbm_nms_boxes": (_c_int, [_vp, _vp, _c_int, _c_float, _c_float, _c_float, _c_int, _vp, _ip]),
    "sbm_set_tuning": (_c_int, [_vp, _c_int, _c_int, _c_int, _c_int, _c_int]),
    "sbm_get_tuning": (_c_int, [_vp, _ip, _ip, _ip, _ip, _ip]),
    "sbm_autotune": (_c_int, [_vp, _vp, _c_float, _cpp, _c_int, _c_int, _ip, _ip, _ip, _ip, _ip, _vp, _c_int, _ip]),
    "sbm_autotune_row": (_c_int, [_vp, _c_int, _ip, _vp, _fp, _ip]),
    "sbm_autotune_images": (_c_int, [_vp, _vp, _c_int, _c_float, _cpp, _c_int, _c_int]),
    "sbm_set_tuning_param": (_c_int, [_vp, _cp, _c_int]),
    "sbm_get_tuning_param": (_c_int, [_vp, _cp, _ip]),
    "sbm_set_timing": (_c_int, [_vp, _c_int]),
    "sbm_set_coarse_mode": (_c_int, [_vp, _c_int]),
    "sbm_get_coarse_mode": (_c_int, [_vp, _ip]),
    "sbm_last_coarse_path": (_c_int, [_vp, _ip, _ip]),
    "sbm_last_coarse_entries": (_c_int, [_vp, ctypes.POINTER(ctypes.c_longlong)]),
    "sbm_last_refine_features": (_c_int, [_vp, ctypes.POINTER(ctypes.c_longlong)]),
    "sbm_last_device_ms": (_c_int, [_vp, _fp]),
    "sbm_last_timing": (_c_int, [_vp, _fp, _fp, _fp, _ip, _ip]),
    "sbm_last_kernel_time": (_c_int, [_vp, _c_int, _cp, _c_int, _fp, _ip]),
    "sbm_coarse_gather_bytes": (_c_int, [_vp, _c_int, _c_int, _cpp, _c_int, _llp]),
    "sbm_device_bytes": (_c_int, [_vp, _llp]),
    "sbm_copy_probe": (_c_int, [_vp, ctypes.c_longlong]),
    "sbm_prepare_image": (_c_int, [_vp, _vp, _c_int, _c_int, _c_int, _c_int, _vp]),
    "sbm_read_level": (_c_int, [_vp, _c_int, _c_int, _vp]),
    "sbm_similarity_map": (_c_int, [_vp, _cp, _c_int, _vp]),
    "sbm_similarity_local": (_c_int, [_vp, _cp, _c_int, _c_int, _c_int, _c_int, _vp]),
    "sbm_gradient": (_c_int, [_vp, _vp, _c_int, _c_int, _c_int, _c_int, _vp, _vp, _vp]),
    "sbm_bin_table": (_c_int, [_vp, _c_int, _c_int, _vp]),
    "sbm_division_check": (_c_int, [_vp, _vp]),
}

_lib = None


class SbmError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[{code}] {msg}")
        self.code = code


def header_symbols():
    """Names of the entry points include/sbm_hip.h declares."""
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(sbm_\w+)\s*\(", text, re.M)))


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: build it with `python -m shape_based_matching_amd.build` "
                "(the matching path has no CPU fallback)")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc):
    if rc != SBM_OK:
        msg = lib().sbm_last_error()
        raise SbmError(rc, msg.decode() if msg else "")
    return rc
