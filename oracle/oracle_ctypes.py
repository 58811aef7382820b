"""ctypes binding of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker or the CPU baseline.  Never used
by the product path.
"""
from __future__ import annotations

import ctypes as C
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(os.path.dirname(HERE), "3deye_amd"))
import kfx  # noqa: E402  (structure definitions only)

LIB_PATH = os.path.join(HERE, "liboracle.so")

_P = C.c_void_p
_I = C.c_int
RB = C.POINTER(kfx.RoundBuffer)
IN = C.POINTER(kfx.Intrinsics)
M3 = C.POINTER(kfx.Matrix3f)
P3 = C.POINTER(kfx.Point3f)
PD = C.POINTER(C.c_double)

_PROTOS = {
    "orc_set_threads": (None, [_I]),
    "orc_get_threads": (_I, []),
    "orc_cyclic_address": (C.c_uint32, [_I, _I, _I, RB]),
    "orc_tsdf_decode": (C.c_float, [C.c_int16]),
    "orc_tsdf_encode": (C.c_int16, [C.c_float]),
    "orc_scale_depth_f32": (None, [_P, _P, IN]),
    "orc_scale_depth_f64": (None, [_P, _P, IN]),
    "orc_integrate_f32": (C.c_uint64, [_P, _P, IN, M3, P3, C.c_float, P3, RB]),
    "orc_integrate_f64": (C.c_uint64, [_P, _P, IN, PD, PD, C.c_double, PD, RB]),
    "orc_integrate_colors": (C.c_uint64, [_P, _P, _P, IN, M3, P3, C.c_float, _I, P3, RB]),
    "orc_raycast_f32": (C.c_uint64, [_P, _P, _P, RB, P3, M3, C.c_float, IN, P3]),
    "orc_interpolate_trilinear_f32": (C.c_float, [_P, C.c_float, C.c_float, C.c_float, P3, RB]),
    "orc_bilateral": (None, [_P, _P, _I, _I, C.c_float, C.c_float, C.c_uint32]),
    "orc_pyr_down": (None, [_P, _I, _I, _P, C.c_float]),
    "orc_vertex_map": (None, [_P, _P, IN]),
    "orc_normal_map_f64": (None, [_P, _P, _I, _I]),
    "orc_transform_maps": (None, [_P, _P, M3, P3, _I, _I, _P, _P]),
    "orc_resize_map": (None, [_P, _I, _I, _P, _I]),
    "orc_icp_sums": (_I, [_P, _P, _P, _P, PD, PD, PD, PD, IN, C.c_double, C.c_double, _I, PD]),
    "orc_estimate_combined_f32": (None, [_P, _P, _P, _P, _P, _P, _P, _P, IN, C.c_float, C.c_float, _P]),
    "orc_icp_solve_update": (_I, [PD, PD, PD, PD]),
    "orc_rgbd_prepare": (None, [_P, _P, _I, _I, _I, _P]),
    "orc_rgbd_corr": (_I, [_P, _P, _P, _P, _P, IN, _P, _P, C.c_float, C.c_float, _P]),
    "orc_rgbd_sums": (_I, [_P, _P, _P, _P, _P, IN, _P, PD, PD, PD, PD]),
    "orc_rgbd_twist_exp": (None, [PD, PD, PD]),
    "orc_rgbd_estimate": (_I, [_P, _P, _P, _P, IN, _I, C.POINTER(_I), C.c_float, C.c_float, PD, PD, PD, PD,
                               C.POINTER(_I)]),
    "orc_dn_param_count": (_I, []),
    "orc_dn_forward": (None, [_P, _I, _I, _P, _I, _P, _P]),
    "orc_dn_loss_grad": (C.c_double, [_P, _P, _I, _I, _P, _I, C.c_double, _P, C.POINTER(_I)]),
    "orc_clean_slice": (None, [C.POINTER(kfx.Point3i), C.POINTER(kfx.Point3i), RB, _P, _I]),
    "orc_extract_slice": (C.c_uint64, [C.POINTER(kfx.Point3i), C.POINTER(kfx.Point3i), RB, _P, _P, _P, _P,
                                       C.c_uint64]),
}


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"liboracle.so not built at {LIB_PATH}; run make -C oracle")
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in _PROTOS.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib


lib = _load()


def register(name, restype, argtypes):
    f = getattr(lib, name)
    f.restype = restype
    f.argtypes = argtypes
    return f


def ptr(a):
    return a.ctypes.data_as(C.c_void_p)


def set_threads(n):
    lib.orc_set_threads(int(n))


def scale_depth_f32(depth, intr):
    out = np.zeros(depth.shape, np.float32)
    lib.orc_scale_depth_f32(ptr(np.ascontiguousarray(depth)), ptr(out), C.byref(intr))
    return out


def scale_depth_f64(depth, intr):
    out = np.zeros(depth.shape, np.float64)
    lib.orc_scale_depth_f64(ptr(np.ascontiguousarray(depth)), ptr(out), C.byref(intr))
    return out


def integrate_f32(vol, depth, intr, R_inv, t, trunc, rb):
    """In-place update of `vol` (structured TSDF array); returns N_upd."""
    scaled = scale_depth_f32(depth, intr)
    cs = kfx.cell_size(rb)
    return lib.orc_integrate_f32(ptr(scaled), ptr(vol), C.byref(intr), C.byref(kfx.matrix3f(R_inv)),
                                 C.byref(kfx.point3f(t)), float(trunc), C.byref(cs), C.byref(rb))


def integrate_f64(vol, depth, intr, R_inv, t, trunc, rb):
    scaled = scale_depth_f64(depth, intr)
    R = np.ascontiguousarray(np.asarray(R_inv, np.float64).reshape(9))
    tt = np.ascontiguousarray(np.asarray(t, np.float64).reshape(3))
    cs = np.array([rb.volume_size.x / rb.voxels_size.x, rb.volume_size.y / rb.voxels_size.y,
                   rb.volume_size.z / rb.voxels_size.z], np.float64)
    return lib.orc_integrate_f64(ptr(scaled), ptr(vol), C.byref(intr), R.ctypes.data_as(PD), tt.ctypes.data_as(PD),
                                 float(trunc), cs.ctypes.data_as(PD), C.byref(rb))


def integrate_colors(cvol, rgb, vmap, intr, R_inv, t, trunc, max_weight, rb):
    """In-place update of `cvol` (COLOR_DTYPE array of buffer_length) from
    `rgb` (rows x cols COLOR_DTYPE) and the raycast `vmap` (rows x cols x 3
    float32, volume coordinates); returns the number of voxels touched."""
    cs = kfx.cell_size(rb)
    rgb = np.ascontiguousarray(rgb)
    vmap = np.ascontiguousarray(vmap, np.float32)
    assert cvol.dtype == COLOR_DTYPE and cvol.size == rb.buffer_length
    assert rgb.dtype == COLOR_DTYPE and rgb.size == intr.rows * intr.cols and vmap.size == 3 * rgb.size
    return lib.orc_integrate_colors(ptr(cvol), ptr(rgb), ptr(vmap), C.byref(intr), C.byref(kfx.matrix3f(R_inv)),
                                    C.byref(kfx.point3f(t)), float(trunc), int(max_weight), C.byref(cs),
                                    C.byref(rb))


def raycast_f32(vol, rb, t, R, trunc, intr, want_samples=False):
    vmap = np.zeros((intr.rows, intr.cols, 3), np.float32)
    nmap = np.zeros((intr.rows, intr.cols, 3), np.float32)
    cs = kfx.cell_size(rb)
    samples = lib.orc_raycast_f32(ptr(vol), ptr(vmap), ptr(nmap), C.byref(rb), C.byref(kfx.point3f(t)),
                                  C.byref(kfx.matrix3f(R)), float(trunc), C.byref(intr), C.byref(cs))
    return (vmap, nmap, samples) if want_samples else (vmap, nmap)


NO_MAX_DISTANCE = 0xFFFFFFFF


def bilateral(depth, sigma_space=4.5, sigma_color=30.0, max_distance=NO_MAX_DISTANCE):
    depth = np.ascontiguousarray(depth, np.uint16)
    out = np.zeros_like(depth)
    lib.orc_bilateral(ptr(depth), ptr(out), depth.shape[1], depth.shape[0], sigma_space, sigma_color, max_distance)
    return out


def pyr_down(depth, sigma_color=30.0):
    depth = np.ascontiguousarray(depth, np.uint16)
    out = np.zeros((depth.shape[0] // 2, depth.shape[1] // 2), np.uint16)
    lib.orc_pyr_down(ptr(depth), depth.shape[1], depth.shape[0], ptr(out), sigma_color)
    return out


def vertex_map(depth, intr):
    depth = np.ascontiguousarray(depth, np.uint16)
    out = np.zeros(depth.shape + (3,), np.float32)
    lib.orc_vertex_map(ptr(depth), ptr(out), C.byref(intr))
    return out


def normal_map_f64(vmap):
    vmap = np.ascontiguousarray(vmap, np.float32)
    out = np.zeros_like(vmap)
    lib.orc_normal_map_f64(ptr(vmap), ptr(out), vmap.shape[1], vmap.shape[0])
    return out


def transform_maps(vmap, nmap, R, t):
    vmap = np.ascontiguousarray(vmap, np.float32)
    nmap = np.ascontiguousarray(nmap, np.float32)
    vd, nd = np.zeros_like(vmap), np.zeros_like(nmap)
    lib.orc_transform_maps(ptr(vmap), ptr(nmap), C.byref(kfx.matrix3f(R)), C.byref(kfx.point3f(t)),
                           vmap.shape[1], vmap.shape[0], ptr(vd), ptr(nd))
    return vd, nd


def resize_map(m, normalize):
    m = np.ascontiguousarray(m, np.float32)
    out = np.zeros((m.shape[0] // 2, m.shape[1] // 2, 3), np.float32)
    lib.orc_resize_map(ptr(m), m.shape[1], m.shape[0], ptr(out), int(normalize))
    return out


def _d(a, n):
    a = np.ascontiguousarray(np.asarray(a, np.float64).reshape(n))
    return a, a.ctypes.data_as(PD)


def icp_sums(vc, nc, vp, np_, Rc, tc, Rprev_inv, tprev, intr, dthr, athr, precision=0):
    keep = [_d(Rc, 9), _d(tc, 3), _d(Rprev_inv, 9), _d(tprev, 3)]
    out = np.zeros(27, np.float64)
    n = lib.orc_icp_sums(ptr(vc), ptr(nc), ptr(vp), ptr(np_), keep[0][1], keep[1][1], keep[2][1], keep[3][1],
                         C.byref(intr), float(dthr), float(athr), int(precision), out.ctypes.data_as(PD))
    return out, n


def estimate_combined_f32(vc, nc, vp, np_, Rc, tc, Rprev_inv, tprev, intr, dthr, athr):
    f = [np.ascontiguousarray(np.asarray(a, np.float32).reshape(-1)) for a in (Rc, tc, Rprev_inv, tprev)]
    out = np.zeros(27, np.float32)
    lib.orc_estimate_combined_f32(ptr(vc), ptr(nc), ptr(vp), ptr(np_), ptr(f[0]), ptr(f[1]), ptr(f[2]), ptr(f[3]),
                                  C.byref(intr), float(dthr), float(athr), ptr(out))
    return out


def icp_solve_update(sums, R, t):
    s, sp = _d(sums, 27)
    R = np.array(R, np.float64).reshape(9)
    t = np.array(t, np.float64).reshape(3)
    x = np.zeros(6, np.float64)
    rc = lib.orc_icp_solve_update(sp, R.ctypes.data_as(PD), t.ctypes.data_as(PD), x.ctypes.data_as(PD))
    return rc, R.reshape(3, 3), t, x


def unpack_sums(s):
    """27 sums -> (A, b) as opencl_wrappers.cpp:426-437."""
    A = np.zeros((6, 6))
    b = np.zeros(6)
    k = 0
    for i in range(6):
        for j in range(i, 7):
            if j == 6:
                b[i] = s[k]
            else:
                A[i, j] = A[j, i] = s[k]
            k += 1
    return A, b


COLOR_DTYPE = np.dtype([("r", "u1"), ("g", "u1"), ("b", "u1"), ("w", "u1")])


def clean_slice(vol, mn, mx, rb):
    lib.orc_clean_slice(C.byref(kfx.point3i(mn)), C.byref(kfx.point3i(mx)), C.byref(rb), ptr(vol), vol.dtype.itemsize)


def extract_slice(vol, col, mn, mx, rb):
    """-> (points float32 [n,4], colors uint8 [n,4])."""
    n = lib.orc_extract_slice(C.byref(kfx.point3i(mn)), C.byref(kfx.point3i(mx)), C.byref(rb), ptr(vol),
                              ptr(col) if col is not None else None, None, None, 0)
    pts = np.zeros((n, 4), np.float32)
    cols = np.zeros((n, 4), np.uint8)
    lib.orc_extract_slice(C.byref(kfx.point3i(mn)), C.byref(kfx.point3i(mx)), C.byref(rb), ptr(vol),
                          ptr(col) if col is not None else None, ptr(pts), ptr(cols), n)
    return pts, cols


def dn_forward(depth, params, bf16=True, hidden=False):
    depth = np.ascontiguousarray(depth, np.uint16)
    H, W = depth.shape
    params = np.ascontiguousarray(params, np.float32)
    y = np.zeros((H, W), np.float64)
    hid = np.zeros((3, H, W, 16), np.float64) if hidden else None
    lib.orc_dn_forward(ptr(depth), W, H, ptr(params), int(bf16), ptr(y), ptr(hid) if hidden else None)
    return (y, hid) if hidden else y


def dn_loss_grad(depth, target, params, bf16=True, outlier=0.05):
    depth = np.ascontiguousarray(depth, np.uint16)
    H, W = depth.shape
    target = np.ascontiguousarray(target, np.float32)
    params = np.ascontiguousarray(params, np.float32)
    g = np.zeros(params.size, np.float64)
    n = C.c_int()
    loss = lib.orc_dn_loss_grad(ptr(depth), ptr(target), W, H, ptr(params), int(bf16), float(outlier), ptr(g),
                                C.byref(n))
    return loss, g, n.value


def rgbd_prepare(rgb, depth, levels):
    """Level arrays of one frame: a list of (I, D, dx, dy) per level."""
    rgb = np.ascontiguousarray(rgb, np.uint8)
    depth = np.ascontiguousarray(depth, np.uint16)
    H, W = depth.shape
    sizes = [(H >> l, W >> l) for l in range(levels)]
    out = np.zeros(sum(4 * h * w for h, w in sizes), np.float32)
    lib.orc_rgbd_prepare(ptr(rgb), ptr(depth), W, H, levels, ptr(out))
    res, off = [], 0
    for h, w in sizes:
        arrs = [out[off + k * h * w: off + (k + 1) * h * w].reshape(h, w) for k in range(4)]
        res.append(tuple(arrs))
        off += 4 * h * w
    return res


def rgbd_corr(prev, curr, intr, Rinv, tinv, min_gradient, max_depth_diff):
    """prev, curr: (I, D, dx, dy) of one level; returns the pair map (-1 none) and the pair count."""
    H, W = prev[1].shape
    corr = np.zeros(H * W, np.int32)
    Ri = np.ascontiguousarray(Rinv, np.float32).reshape(9)
    ti = np.ascontiguousarray(tinv, np.float32).reshape(3)
    n = lib.orc_rgbd_corr(ptr(prev[0]), ptr(prev[1]), ptr(curr[1]), ptr(curr[2]), ptr(curr[3]), C.byref(intr),
                          ptr(Ri), ptr(ti), float(min_gradient) ** 2, float(max_depth_diff), ptr(corr))
    return corr.reshape(H, W), n


def rgbd_sums(prev, curr, intr, corr, R, t):
    Rd = np.ascontiguousarray(R, np.float64).reshape(9)
    td = np.ascontiguousarray(t, np.float64).reshape(3)
    sigma = C.c_double()
    s = np.zeros(27, np.float64)
    c = np.ascontiguousarray(corr, np.int32)
    n = lib.orc_rgbd_sums(ptr(prev[0]), ptr(prev[1]), ptr(curr[0]), ptr(curr[2]), ptr(curr[3]), C.byref(intr),
                          ptr(c), Rd.ctypes.data_as(PD), td.ctypes.data_as(PD), C.byref(sigma),
                          s.ctypes.data_as(PD))
    return s, sigma.value, n


def rgbd_twist_exp(ksi):
    k = np.ascontiguousarray(ksi, np.float64)
    R = np.zeros(9)
    t = np.zeros(3)
    lib.orc_rgbd_twist_exp(k.ctypes.data_as(PD), R.ctypes.data_as(PD), t.ctypes.data_as(PD))
    return R.reshape(3, 3), t


def rgbd_estimate(rgb_prev, depth_prev, rgb_curr, depth_curr, intr, levels, iterations, min_gradient,
                  max_depth_diff, R0, t0):
    it = (C.c_int * 4)(*(list(iterations) + [0] * (4 - len(iterations))))
    R0d = np.ascontiguousarray(R0, np.float64).reshape(9)
    t0d = np.ascontiguousarray(t0, np.float64).reshape(3)
    R = np.zeros(9)
    t = np.zeros(3)
    solves = C.c_int()
    arrs = [np.ascontiguousarray(a) for a in (rgb_prev, depth_prev, rgb_curr, depth_curr)]
    ok = lib.orc_rgbd_estimate(ptr(arrs[0]), ptr(arrs[1]), ptr(arrs[2]), ptr(arrs[3]), C.byref(intr), levels, it,
                               float(min_gradient), float(max_depth_diff), R0d.ctypes.data_as(PD),
                               t0d.ctypes.data_as(PD), R.ctypes.data_as(PD), t.ctypes.data_as(PD), C.byref(solves))
    return ok, R.reshape(3, 3), t, solves.value
