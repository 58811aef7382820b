"""ctypes binding of libkfx (include/kfx.h) for Python callers: the tests,
bench.py and __graft_entry__.

The binding is a thin mirror of the C ABI: ctypes structures with the exact
layouts of include/kfx_types.h, a status check that raises KfxError with
kfx_last_error(), and a small device-buffer helper that moves numpy arrays in
and out of HIP memory through kfx_memcpy_*.  There is no fallback: if
libkfx.so is missing the import of this module fails.
"""
from __future__ import annotations

import ctypes as C
import os
import re

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libkfx.so")
HEADER_DIR = os.path.join(os.path.dirname(HERE), "include")


class KfxError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"kfx status {code}: {msg}")
        self.code = code


# --------------------------------------------------------------------------
# Structures (include/kfx_types.h)
# --------------------------------------------------------------------------
class Point3f(C.Structure):
    _fields_ = [("x", C.c_float), ("y", C.c_float), ("z", C.c_float)]


class Point3i(C.Structure):
    _fields_ = [("x", C.c_int32), ("y", C.c_int32), ("z", C.c_int32)]


class Matrix3f(C.Structure):
    _fields_ = [("data", Point3f * 3)]


class Intrinsics(C.Structure):
    _fields_ = [("fx", C.c_float), ("fy", C.c_float), ("cx", C.c_float), ("cy", C.c_float),
                ("cols", C.c_int32), ("rows", C.c_int32)]

    def level(self, i):
        """intrinsics::operator()(i) (common/src/intrinsics.cpp:7-18)."""
        d = float(1 << i)
        return Intrinsics(self.fx / d, self.fy / d, self.cx / d, self.cy / d, self.cols >> i, self.rows >> i)


class RoundBuffer(C.Structure):
    _fields_ = [("start_index", C.c_int32), ("end_index", C.c_int32), ("buffer_length", C.c_int32),
                ("origin_GRID", Point3i), ("origin_GRID_global", Point3f), ("origin_metric", Point3f),
                ("volume_size", Point3f), ("voxels_size", Point3i)]


class SynthObject(C.Structure):
    _fields_ = [("shape", C.c_int32), ("reserved", C.c_int32), ("center", C.c_double * 3),
                ("half", C.c_double * 3)]


class SynthScene(C.Structure):
    _fields_ = [("ground_y", C.c_double), ("n_objects", C.c_int32), ("reserved", C.c_int32),
                ("objects", SynthObject * 64)]


class Pose(C.Structure):
    _fields_ = [("R", C.c_double * 9), ("t", C.c_double * 3)]

    def matrices(self):
        return np.array(self.R[:], dtype=np.float64).reshape(3, 3), np.array(self.t[:], dtype=np.float64)


def pose_from(R, t):
    p = Pose()
    p.R[:] = [float(v) for v in np.asarray(R, np.float64).reshape(9)]
    p.t[:] = [float(v) for v in np.asarray(t, np.float64).reshape(3)]
    return p


class IcpResult(C.Structure):
    _fields_ = [("pose", Pose), ("A", C.c_double * 36), ("b", C.c_double * 6), ("x", C.c_double * 6),
                ("status", C.c_int32), ("singular_count", C.c_int32), ("iterations", C.c_int32),
                ("converged_tag", C.c_int32)]


class RgbdParams(C.Structure):
    _fields_ = [("max_depth_diff", C.c_float), ("min_gradient", C.c_float), ("levels", C.c_int32),
                ("iterations", C.c_int32 * 4)]


class RgbdResult(C.Structure):
    _fields_ = [("pose", Pose), ("A", C.c_double * 36), ("b", C.c_double * 6), ("sigma", C.c_double),
                ("correspondences", C.c_int32), ("solves", C.c_int32), ("ok", C.c_int32), ("reserved", C.c_int32)]


def rgbd_params(max_depth_diff=0.07, min_gradient=10.0, levels=4, iterations=(7, 7, 7, 10)):
    """Defaults of the OpenCV RGBD odometry the reference follows."""
    it = list(iterations) + [0] * (4 - len(iterations))
    return RgbdParams(max_depth_diff, min_gradient, levels, (C.c_int32 * 4)(*it))


class Xyzi(C.Structure):
    _fields_ = [("x", C.c_float), ("y", C.c_float), ("z", C.c_float), ("w", C.c_float)]


class IcpLevel(C.Structure):
    _fields_ = [("vmap_current", C.c_void_p), ("nmap_current", C.c_void_p), ("vmap_previous", C.c_void_p),
                ("nmap_previous", C.c_void_p), ("intr", Intrinsics), ("iterations", C.c_int32)]


class PipelineConfig(C.Structure):
    _fields_ = [("intr", Intrinsics), ("voxels", Point3i), ("volume_size", Point3f), ("truncation", C.c_float),
                ("sigma_space", C.c_float), ("sigma_color", C.c_float), ("distance_threshold", C.c_float),
                ("angle_threshold", C.c_float), ("icp_iterations", C.c_int32 * 3), ("mode", C.c_int32),
                ("use_graph", C.c_int32), ("denoiser", C.c_int32), ("fallback_rms_mm", C.c_float),
                ("shift_threshold", C.c_float), ("integrate_threshold", C.c_float),
                ("converge_rotation", C.c_double), ("converge_translation", C.c_double),
                ("extract_on_shift", C.c_int32), ("icp_sum_mode", C.c_int32), ("shift_target_distance", C.c_double),
                ("colors", C.c_int32), ("color_max_weight", C.c_int32), ("extract_records", C.c_int32),
                ("icp_launch", C.c_int32)]


class FrameInfo(C.Structure):
    _fields_ = [("pose_world", Pose), ("pose_volume", Pose), ("icp_iterations", C.c_int32), ("icp_singular", C.c_int32),
                ("integrated", C.c_int32), ("shifted", C.c_int32), ("extracted_points", C.c_uint64),
                ("fallback", C.c_int32), ("denoise_rms_mm", C.c_float)]


class PipelineView(C.Structure):
    _fields_ = [("volume", C.c_void_p), ("round_buffer", RoundBuffer), ("cell_size", Point3f),
                ("origin_world", C.c_double * 3), ("depth_raw", C.c_void_p), ("depth_filtered", C.c_void_p),
                ("vmap_model", C.c_void_p), ("nmap_model", C.c_void_p), ("vmap_current", C.c_void_p),
                ("nmap_current", C.c_void_p), ("pose_volume", Pose), ("frames", C.c_int64), ("shifts", C.c_int32),
                ("fallbacks", C.c_int32), ("extracted_points", C.c_uint64), ("last_points", C.c_void_p),
                ("last_colors", C.c_void_p), ("color_volume", C.c_void_p), ("rgb", C.c_void_p),
                ("last_records", C.c_void_p), ("brick_flags", C.c_void_p),
                ("vmap_model_levels", C.c_void_p * 3), ("nmap_model_levels", C.c_void_p * 3),
                ("depth_scaled", C.c_void_p), ("depth_bound", C.c_void_p)]


PLAIN, FUSED = 0, 1
ICP_SUM_FP32_KAHAN, ICP_SUM_FP32, ICP_SUM_FP64 = 0, 1, 2
ICP_LAUNCH_FRAME, ICP_LAUNCH_TWO, ICP_LAUNCH_ONE = 0, 1, 2

assert C.sizeof(Intrinsics) == 24 and C.sizeof(RoundBuffer) == 72 and C.sizeof(Matrix3f) == 36

TSDF_DTYPE = np.dtype([("value", "<i2"), ("weight", "<i2")])
XYZ_DTYPE = np.dtype([("x", "<f4"), ("y", "<f4"), ("z", "<f4")])
COLOR_DTYPE = np.dtype([("r", "u1"), ("g", "u1"), ("b", "u1"), ("w", "u1")])
SLICE_RECORD_DTYPE = np.dtype([("x", "<i4"), ("y", "<i4"), ("z", "<i4"), ("tsdf", "<i2"), ("weight", "<i2")])


def point3f(v):
    return Point3f(float(v[0]), float(v[1]), float(v[2]))


def point3i(v):
    return Point3i(int(v[0]), int(v[1]), int(v[2]))


def matrix3f(m):
    m = np.asarray(m, dtype=np.float64).reshape(3, 3)
    return Matrix3f((Point3f * 3)(*[Point3f(*[float(c) for c in m[r]]) for r in range(3)]))


def round_buffer(voxels, volume_m, origin=(0, 0, 0)):
    """round_buffer(point_3i voxels_size, point_3f metric_size)
    (common/src/round_buffer.cpp:65-74): start_index 0, buffer_length Nx*Ny*Nz."""
    n = [int(v) for v in (voxels if hasattr(voxels, "__len__") else (voxels,) * 3)]
    s = [float(v) for v in (volume_m if hasattr(volume_m, "__len__") else (volume_m,) * 3)]
    rb = RoundBuffer()
    rb.start_index = 0
    rb.end_index = 0
    rb.buffer_length = n[0] * n[1] * n[2]
    rb.origin_GRID = point3i(origin)
    rb.origin_GRID_global = Point3f(float(origin[0]), float(origin[1]), float(origin[2]))
    rb.origin_metric = Point3f(0.0, 0.0, 0.0)
    rb.volume_size = point3f(s)
    rb.voxels_size = point3i(n)
    return rb


def cell_size(rb):
    return Point3f(rb.volume_size.x / rb.voxels_size.x, rb.volume_size.y / rb.voxels_size.y,
                   rb.volume_size.z / rb.voxels_size.z)


# --------------------------------------------------------------------------
# Library loading and prototypes
# --------------------------------------------------------------------------
_P = C.c_void_p
_I = C.c_int
_SZ = C.c_size_t


def header_functions():
    """Names of every function declared in include/*.h."""
    names = []
    for fn in sorted(os.listdir(HEADER_DIR)):
        if not fn.endswith(".h"):
            continue
        text = open(os.path.join(HEADER_DIR, fn)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names += re.findall(r"^\s*(?:const\s+)?[A-Za-z_][\w]*\s*\**\s+\**(kfx_\w+)\s*\(", text, flags=re.M)
    return names


_PROTOS = {
    "kfx_init": (_I, [_I]),
    "kfx_last_error": (C.c_char_p, []),
    "kfx_device_count": (_I, [C.POINTER(_I)]),
    "kfx_synchronize": (_I, []),
    "kfx_get_stream": (_I, [C.POINTER(_P)]),
    "kfx_malloc": (_I, [C.POINTER(_P), _SZ]),
    "kfx_free": (_I, [_P]),
    "kfx_memcpy_htod": (_I, [_P, _P, _SZ]),
    "kfx_memcpy_dtoh": (_I, [_P, _P, _SZ]),
    "kfx_memcpy_dtod": (_I, [_P, _P, _SZ]),
    "kfx_fill_u8": (_I, [_P, C.c_uint8, _SZ]),
    "kfx_copy_kernel": (_I, [_P, _P, _SZ]),
    "kfx_event_create": (_I, [C.POINTER(_P)]),
    "kfx_event_destroy": (_I, [_P]),
    "kfx_event_record": (_I, [_P]),
    "kfx_event_elapsed_ms": (_I, [_P, _P, C.POINTER(C.c_float)]),
    "kfx_cyclic_address": (C.c_uint32, [_I, _I, _I, C.POINTER(RoundBuffer)]),
    "kfx_cyclic_address_batch": (_I, [_P, _P, _SZ, C.POINTER(RoundBuffer)]),
    "kfx_init_tsdf": (_I, [_P, C.POINTER(RoundBuffer)]),
    "kfx_integrate_tsdf": (_I, [_P, _P, _P, C.POINTER(Intrinsics), C.POINTER(Matrix3f), C.POINTER(Point3f),
                                C.c_float, C.POINTER(Point3f), C.POINTER(RoundBuffer)]),
    "kfx_integrate_tsdf_noalloc": (_I, [_P, _P, _P, C.POINTER(Intrinsics), C.POINTER(Matrix3f), C.POINTER(Point3f),
                                        C.c_float, C.POINTER(Point3f), C.POINTER(RoundBuffer)]),
    "kfx_raycast": (_I, [_P, _P, _P, C.POINTER(RoundBuffer), C.POINTER(Point3f), C.POINTER(Matrix3f), C.c_float,
                         C.POINTER(Intrinsics), C.POINTER(Point3f)]),
    "kfx_integrate_tsdf_posed": (_I, [_P, _P, _P, C.POINTER(Intrinsics), _P, _P, C.c_float, C.c_float,
                                      C.POINTER(Point3f), C.POINTER(RoundBuffer), _P]),
    "kfx_raycast_posed": (_I, [_P, _P, _P, C.POINTER(RoundBuffer), _P, C.c_float, C.POINTER(Intrinsics),
                               C.POINTER(Point3f), _P]),
    "kfx_brick_flags_size": (_I, [C.POINTER(RoundBuffer), C.POINTER(_SZ)]),
    "kfx_brick_flags_build": (_I, [_P, C.POINTER(RoundBuffer), _P]),
    "kfx_raycast_bricks": (_I, [_P, _P, _P, C.POINTER(RoundBuffer), C.POINTER(Point3f), C.POINTER(Matrix3f),
                                C.c_float, C.POINTER(Intrinsics), C.POINTER(Point3f), _P]),
    "kfx_raycast_count_samples": (_I, [_P, _P, _P, C.POINTER(RoundBuffer), C.POINTER(Point3f), C.POINTER(Matrix3f),
                                       C.c_float, C.POINTER(Intrinsics), C.POINTER(Point3f),
                                       C.POINTER(C.c_uint64)]),
    "kfx_maps_pyramid": (_I, [_P, _P, _P, C.POINTER(Intrinsics), _I]),
    "kfx_maps_pyramid_from_depth": (_I, [_P, _P, _P, _P, C.POINTER(Intrinsics), _I, C.c_float]),
    "kfx_resize_pyramid": (_I, [_P, _P, _I, _I, _P, _P, _P, _P]),
    "kfx_integrate_colors": (_I, [_P, _P, _P, C.POINTER(Intrinsics), C.POINTER(Matrix3f), C.POINTER(Point3f),
                                  C.c_float, _I, C.POINTER(Point3f), C.POINTER(RoundBuffer)]),
    "kfx_integrate_colors_posed": (_I, [_P, _P, _P, C.POINTER(Intrinsics), _P, C.c_float, _I, C.POINTER(Point3f),
                                        C.POINTER(RoundBuffer), _P]),
    "kfx_init_colors": (_I, [_P, C.POINTER(RoundBuffer)]),
    "kfx_profile_enable": (_I, [_I]),
    "kfx_profile_reset": (_I, []),
    "kfx_profile_query": (_I, [C.c_char_p, C.POINTER(C.c_double), C.POINTER(C.c_uint64)]),
    "kfx_integrate_count_updates": (_I, [_P, _P, C.POINTER(Intrinsics), C.POINTER(Matrix3f), C.POINTER(Point3f),
                                         C.c_float, C.POINTER(Point3f), C.POINTER(RoundBuffer),
                                         C.POINTER(C.c_uint64)]),
    "kfx_bilateral_filter": (_I, [_P, _P, _I, _I, C.c_float, C.c_float, C.c_uint32]),
    "kfx_pyr_down": (_I, [_P, _I, _I, _P, C.c_float]),
    "kfx_vertex_map": (_I, [_P, _P, C.POINTER(Intrinsics)]),
    "kfx_normal_map": (_I, [_P, _P, _I, _I]),
    "kfx_transform_maps": (_I, [_P, _P, C.POINTER(Matrix3f), C.POINTER(Point3f), _I, _I, _P, _P]),
    "kfx_resize_map": (_I, [_P, _I, _I, _P, _I]),
    "kfx_estimate_combined": (_I, [C.POINTER(Matrix3f), C.POINTER(Point3f), _P, _P, C.POINTER(Matrix3f),
                                   C.POINTER(Point3f), C.POINTER(Intrinsics), _P, _P, C.c_float, C.c_float, _P, _I,
                                   _P, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "kfx_clean_buff": (_I, [_P, _I]),
    "kfx_icp_create": (_I, [_I, C.POINTER(_P)]),
    "kfx_icp_set_sum_mode": (_I, [_P, _I]),
    "kfx_icp_destroy": (_I, [_P]),
    "kfx_icp_reset": (_I, [_P, C.POINTER(Pose)]),
    "kfx_icp_iterate": (_I, [_P, _P, _P, _P, _P, C.POINTER(Matrix3f), C.POINTER(Point3f), C.POINTER(Intrinsics),
                             C.c_float, C.c_float, _I, C.c_double, C.c_double]),
    "kfx_check_reciprocal": (_I, [_I, _I, C.POINTER(C.c_uint64)]),
    "kfx_icp_read": (_I, [_P, C.POINTER(IcpResult)]),
    "kfx_icp_publish_next": (_I, [_P]),
    "kfx_icp_set_launch_mode": (_I, [_P, _I]),
    "kfx_icp_read_published": (_I, [_P, C.POINTER(IcpResult)]),
    "kfx_rgbd_create": (_I, [_I, _I, _I, C.POINTER(_P)]),
    "kfx_rgbd_destroy": (_I, [_P]),
    "kfx_rgbd_set_frame": (_I, [_P, _I, _P, _P]),
    "kfx_rgbd_swap": (_I, [_P]),
    "kfx_rgbd_download_level": (_I, [_P, _I, _I, _P, _P, _P, _P]),
    "kfx_rgbd_estimate": (_I, [_P, C.POINTER(Intrinsics), C.POINTER(RgbdParams), C.POINTER(Pose),
                               C.POINTER(RgbdResult)]),
    "kfx_icp_begin_frame": (_I, [_P]),
    "kfx_icp_frame": (_I, [_P, C.POINTER(IcpLevel), _I, C.c_float, C.c_float, _I, C.c_double, C.c_double]),
    "kfx_icp_frame_resident": (_I, [_P, C.POINTER(_I)]),
    "kfx_icp_solve_sums": (_I, [_P, _P]),
    "kfx_icp_translate": (_I, [_P, C.c_double, C.c_double, C.c_double]),
    "kfx_icp_pose_device": (_I, [_P, C.POINTER(_P), C.POINTER(_P)]),
    "kfx_determine_extract_step": (_I, [_I, C.POINTER(_I), C.POINTER(_I)]),
    "kfx_clean_tsdf_slice": (_I, [C.POINTER(Point3i), C.POINTER(Point3i), C.POINTER(RoundBuffer), _P]),
    "kfx_clean_color_slice": (_I, [C.POINTER(Point3i), C.POINTER(Point3i), C.POINTER(RoundBuffer), _P]),
    "kfx_slicer_create": (_I, [C.POINTER(RoundBuffer), C.POINTER(_P)]),
    "kfx_slicer_destroy": (_I, [_P]),
    "kfx_extract_slice": (_I, [_P, C.POINTER(Point3i), C.POINTER(Point3i), C.POINTER(RoundBuffer), _P, _P, _P, _P,
                               C.c_uint64, C.POINTER(C.c_uint64)]),
    "kfx_extract_slice_records": (_I, [_P, C.POINTER(Point3i), C.POINTER(Point3i), C.POINTER(RoundBuffer), _P, _P,
                                       _P, _P, C.c_uint64, C.POINTER(C.c_uint64)]),
    "kfx_pipeline_default_config": (_I, [C.POINTER(PipelineConfig)]),
    "kfx_pipeline_create": (_I, [C.POINTER(PipelineConfig), C.POINTER(Pose), C.POINTER(C.c_double), C.POINTER(_P)]),
    "kfx_pipeline_destroy": (_I, [_P]),
    "kfx_pipeline_process": (_I, [_P, _P, C.POINTER(FrameInfo)]),
    "kfx_pipeline_process_rgbd": (_I, [_P, _P, _P, C.POINTER(FrameInfo)]),
    "kfx_pipeline_state": (_I, [_P, C.POINTER(PipelineView)]),
    "kfx_pipeline_set_denoiser": (_I, [_P, _P]),
    "kfx_denoiser_param_count": (_I, []),
    "kfx_denoiser_create": (_I, [_I, _I, C.POINTER(_P)]),
    "kfx_denoiser_destroy": (_I, [_P]),
    "kfx_denoiser_set_params": (_I, [_P, _P]),
    "kfx_denoiser_get_params": (_I, [_P, _P]),
    "kfx_denoiser_forward": (_I, [_P, _P, _P, _P]),
    "kfx_denoiser_forward_layers": (_I, [_P, _P, _P, _P]),
    "kfx_denoiser_accumulate": (_I, [_P, _P, _P, C.c_float, C.POINTER(C.c_double), C.POINTER(C.c_uint64)]),
    "kfx_denoiser_get_grad": (_I, [_P, _P]),
    "kfx_denoiser_zero_grad": (_I, [_P]),
    "kfx_denoiser_adam_step": (_I, [_P, C.c_float, _I]),
    "kfx_model_depth": (_I, [_P, C.POINTER(Pose), _I, _I, _P]),
    "kfx_depth_consistency": (_I, [_P, _P, _P, C.POINTER(Intrinsics), C.c_float, _P]),
    "kfx_trajectory_write": (_I, [C.c_char_p, C.POINTER(Pose), _I]),
    "kfx_trajectory_read": (_I, [C.c_char_p, C.POINTER(Pose), _I, C.POINTER(_I)]),
    "kfx_trajectory_errors": (_I, [C.POINTER(Pose), C.POINTER(Pose), _I, C.POINTER(C.c_double),
                                   C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "kfx_splitmix64": (C.c_uint64, [C.c_uint64]),
    "kfx_synth_make_scene": (_I, [C.c_uint64, _I, _I, C.POINTER(C.c_double), C.c_double, C.POINTER(SynthScene)]),
    "kfx_synth_orbit": (_I, [C.POINTER(C.c_double), C.c_double, C.c_double, C.c_double, C.c_double, _I,
                             C.POINTER(Pose)]),
    "kfx_synth_walk": (_I, [C.c_uint64, C.POINTER(C.c_double), C.c_double, C.c_double, _I, C.POINTER(Pose)]),
    "kfx_extract_colors": (_I, [_P, C.POINTER(RoundBuffer), _P, C.c_uint64, C.POINTER(Point3f), _P]),
    "kfx_save_points_text": (_I, [C.c_char_p, _P, C.c_uint64]),
    "kfx_save_colors_text": (_I, [C.c_char_p, _P, C.c_uint64]),
    "kfx_save_pcd": (_I, [C.c_char_p, _P, _P, C.c_uint64, _I]),
    "kfx_save_records": (_I, [C.c_char_p, _P, C.c_uint64]),
    "kfx_synth_render_rgb": (_I, [C.POINTER(SynthScene), C.POINTER(Pose), C.POINTER(Intrinsics), _I, _P]),
    "kfx_synth_render": (_I, [C.POINTER(SynthScene), C.POINTER(Pose), C.POINTER(Intrinsics), C.c_uint64, _I, _I, _I,
                              _P, _P]),
}


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"libkfx.so not built at {LIB_PATH}; run __graft_entry__.build()")
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in _PROTOS.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib


lib = _load()


def register(name, restype, argtypes):
    """Declare the prototype of a further entry point."""
    f = getattr(lib, name)
    f.restype = restype
    f.argtypes = argtypes
    _PROTOS[name] = (restype, argtypes)
    return f


def check(rc):
    if rc != 0:
        raise KfxError(rc, lib.kfx_last_error().decode(errors="replace"))
    return rc


def call(name, *args):
    return check(getattr(lib, name)(*args))


def device_count():
    n = C.c_int(0)
    rc = lib.kfx_device_count(C.byref(n))
    return n.value if rc == 0 else 0


def init(device=0):
    check(lib.kfx_init(device))


def synchronize():
    check(lib.kfx_synchronize())


class Buffer:
    """Device allocation owned by Python (gpu_array_base counterpart)."""

    def __init__(self, nbytes):
        self.nbytes = int(nbytes)
        p = C.c_void_p()
        check(lib.kfx_malloc(C.byref(p), self.nbytes))
        self.ptr = p

    @classmethod
    def from_array(cls, a):
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes)
        b.upload(a)
        return b

    def upload(self, a):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        check(lib.kfx_memcpy_htod(self.ptr, a.ctypes.data_as(C.c_void_p), a.nbytes))

    def download(self, dtype, shape):
        out = np.empty(shape, dtype=dtype)
        assert out.nbytes <= self.nbytes
        check(lib.kfx_memcpy_dtoh(out.ctypes.data_as(C.c_void_p), self.ptr, out.nbytes))
        return out

    def fill(self, value=0):
        check(lib.kfx_fill_u8(self.ptr, value, self.nbytes))

    def free(self):
        if self.ptr is not None and self.ptr.value:
            check(lib.kfx_free(self.ptr))
        self.ptr = None

    def __del__(self):
        try:
            if getattr(self, "ptr", None) is not None and self.ptr.value:
                lib.kfx_free(self.ptr)
        except Exception:
            pass


def download(dev_ptr, dtype, shape):
    """Copy device memory the library owns (e.g. a pipeline view's volume) to
    a new host array."""
    out = np.empty(shape, dtype=dtype)
    check(lib.kfx_memcpy_dtoh(out.ctypes.data_as(C.c_void_p), C.c_void_p(dev_ptr), out.nbytes))
    return out


class Timer:
    """HIP-event timer on the library stream."""

    def __init__(self):
        self.a, self.b = C.c_void_p(), C.c_void_p()
        check(lib.kfx_event_create(C.byref(self.a)))
        check(lib.kfx_event_create(C.byref(self.b)))

    def start(self):
        check(lib.kfx_event_record(self.a))

    def stop(self):
        check(lib.kfx_event_record(self.b))

    def ms(self):
        v = C.c_float()
        check(lib.kfx_event_elapsed_ms(self.a, self.b, C.byref(v)))
        return v.value


def profile_enable(on=True):
    check(lib.kfx_profile_enable(1 if on else 0))


def profile_reset():
    check(lib.kfx_profile_reset())


def profile_query(name):
    """(total milliseconds, launches) of a named kernel since the last reset."""
    ms, n = C.c_double(), C.c_uint64()
    check(lib.kfx_profile_query(name.encode(), C.byref(ms), C.byref(n)))
    return ms.value, n.value


# --------------------------------------------------------------------------
# Synthetic input helpers
# --------------------------------------------------------------------------
def make_scene(seed, n_boxes, n_spheres, region, ground_y):
    sc = SynthScene()
    reg = (C.c_double * 4)(*[float(v) for v in region])
    check(lib.kfx_synth_make_scene(seed, n_boxes, n_spheres, reg, float(ground_y), C.byref(sc)))
    return sc


def orbit(center, radius, dy, angle0, dangle, n):
    poses = (Pose * n)()
    check(lib.kfx_synth_orbit((C.c_double * 3)(*center), radius, dy, angle0, dangle, n, poses))
    return poses


def walk(seed, start, step, pitch, n):
    poses = (Pose * n)()
    check(lib.kfx_synth_walk(seed, (C.c_double * 3)(*start), step, pitch, n, poses))
    return poses


def render(scene, pose, intr, seed, frame, noise=True, threads=0, want_clean=False):
    depth = np.zeros((intr.rows, intr.cols), dtype=np.uint16)
    clean = np.zeros((intr.rows, intr.cols), dtype=np.float32) if want_clean else None
    check(lib.kfx_synth_render(C.byref(scene), C.byref(pose), C.byref(intr), seed, frame, 1 if noise else 0, threads,
                               depth.ctypes.data_as(C.c_void_p),
                               clean.ctypes.data_as(C.c_void_p) if clean is not None else None))
    return (depth, clean) if want_clean else depth


def render_rgb(scene, pose, intr, threads=0):
    """Noise-free colour frame of the view (rows x cols COLOR_DTYPE)."""
    rgb = np.zeros((intr.rows, intr.cols), dtype=COLOR_DTYPE)
    check(lib.kfx_synth_render_rgb(C.byref(scene), C.byref(pose), C.byref(intr), threads,
                                   rgb.ctypes.data_as(C.c_void_p)))
    return rgb


def kinect_intrinsics(cols=640, rows=480, f=525.0):
    return Intrinsics(f, f, (cols - 1) / 2.0, (rows - 1) / 2.0, cols, rows)


class Pipeline:
    """Owner of a kfx_pipeline handle."""

    def __init__(self, cfg, initial_pose_world, volume_origin_world):
        self.h = C.c_void_p()
        check(lib.kfx_pipeline_create(C.byref(cfg), C.byref(initial_pose_world),
                                      (C.c_double * 3)(*[float(v) for v in volume_origin_world]), C.byref(self.h)))
        self.cfg = cfg

    def set_denoiser(self, dn):
        """Attach a Denoiser (or None); the config must have denoiser=1 for
        it to replace the bilateral filter."""
        self.dn = dn
        check(lib.kfx_pipeline_set_denoiser(self.h, dn.h if dn is not None else None))

    def process(self, depth_dev_ptr):
        info = FrameInfo()
        check(lib.kfx_pipeline_process(self.h, depth_dev_ptr, C.byref(info)))
        return info

    def process_rgbd(self, depth_dev_ptr, rgb_dev_ptr):
        info = FrameInfo()
        check(lib.kfx_pipeline_process_rgbd(self.h, depth_dev_ptr, rgb_dev_ptr, C.byref(info)))
        return info

    def view(self):
        v = PipelineView()
        check(lib.kfx_pipeline_state(self.h, C.byref(v)))
        return v

    def close(self):
        if self.h is not None and self.h.value:
            check(lib.kfx_pipeline_destroy(self.h))
        self.h = None

    def __del__(self):
        try:
            if getattr(self, "h", None) is not None and self.h.value:
                lib.kfx_pipeline_destroy(self.h)
        except Exception:
            pass


DENOISER_PARAMS = 5089
DENOISER_WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "denoiser_weights.f32")


class Denoiser:
    """Owner of a kfx_denoiser handle (cols x rows frames)."""

    def __init__(self, cols, rows, params=None):
        self.h = C.c_void_p()
        check(lib.kfx_denoiser_create(cols, rows, C.byref(self.h)))
        if params is not None:
            self.set_params(params)

    @staticmethod
    def load_params(path=DENOISER_WEIGHTS):
        p = np.fromfile(path, dtype="<f4")
        if p.size != DENOISER_PARAMS:
            raise ValueError(f"{path}: {p.size} parameters, expected {DENOISER_PARAMS}")
        return p.astype(np.float32)

    def set_params(self, params):
        p = np.ascontiguousarray(params, dtype=np.float32)
        if p.size != DENOISER_PARAMS:
            raise ValueError(f"{p.size} parameters, expected {DENOISER_PARAMS}")
        check(lib.kfx_denoiser_set_params(self.h, p.ctypes.data_as(C.c_void_p)))

    def get_params(self):
        p = np.zeros(DENOISER_PARAMS, np.float32)
        check(lib.kfx_denoiser_get_params(self.h, p.ctypes.data_as(C.c_void_p)))
        return p

    def close(self):
        if self.h is not None and self.h.value:
            check(lib.kfx_denoiser_destroy(self.h))
        self.h = None

    def __del__(self):
        try:
            if getattr(self, "h", None) is not None and self.h.value:
                lib.kfx_denoiser_destroy(self.h)
        except Exception:
            pass


def default_config(**kw):
    cfg = PipelineConfig()
    check(lib.kfx_pipeline_default_config(C.byref(cfg)))
    for k, v in kw.items():
        if k == "icp_iterations":
            cfg.icp_iterations[:] = list(v)
        else:
            setattr(cfg, k, v)
    return cfg


def trajectory_errors(est, gt):
    """(ATE, RPE translation, RPE rotation) of two pose lists with a common
    first frame (kfx_trajectory_errors)."""
    n = min(len(est), len(gt))
    a, b = (Pose * n)(*est[:n]), (Pose * n)(*gt[:n])
    ate, rt, rr = C.c_double(), C.c_double(), C.c_double()
    check(lib.kfx_trajectory_errors(a, b, n, C.byref(ate), C.byref(rt), C.byref(rr)))
    return ate.value, rt.value, rr.value


def ate_rmse(est, gt):
    """Absolute trajectory error (RMSE of camera positions, metres)."""
    return trajectory_errors(est, gt)[0]
