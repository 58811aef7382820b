"""World-model files (host only): the reference's matrix text format for
extracted slices (common/inc/impl/matrix.hpp:411-476 with the element
formats of common/src/base_points.cpp:278-282 and :570-574), PCD files of the
clouds the reference converts for PCL (common/src/math_utils.cpp:216-328),
and the binary slice-record file.  Each file is read back independently."""
import ctypes as C

import numpy as np


def _points():
    p = np.array([[1.5, -0.25, 3.0, 1.0], [1e-7, 123456789.0, -0.0, 1.0], [511.0, 0.5, 17.125, 2.0]], np.float32)
    c = np.array([[10, 20, 30, 1], [255, 0, 7, 0], [1, 2, 3, 4]], np.uint8)
    return p, c


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


def test_points_text_matches_reference_stream_format(kfx, tmp_path):
    p, _ = _points()
    path = str(tmp_path / "pts.tst")
    kfx.check(kfx.lib.kfx_save_points_text(path.encode(), _ptr(p), len(p)))
    lines = open(path).read().split("\n")
    assert lines[0] == "3 1 1 Point_XYZI"
    # operator<<(XYZI): x 'x' y 'x' z 'x' w, default ostream precision (%g);
    # matrix operator<< adds a space after each element and ends the row.
    want = "".join("x".join("%g" % float(v) for v in row) + " " for row in p)
    assert lines[1] == want
    assert lines[1].startswith("1.5x-0.25x3x1 1e-07x1.23457e+08x-0x1 511x0.5x17.125x2 ")
    assert lines[2:] == [""]


def test_colors_text_matches_reference_stream_format(kfx, tmp_path):
    _, c = _points()
    path = str(tmp_path / "col.tst")
    kfx.check(kfx.lib.kfx_save_colors_text(path.encode(), _ptr(c), len(c)))
    lines = open(path).read().split("\n")
    assert lines[0] == "3 1 1 Point_Color"
    # operator<<(Color) ends with its own space, the matrix adds another.
    assert lines[1] == "10.20.30.1  255.0.7.0  1.2.3.4  "


def _read_pcd(path):
    raw = open(path, "rb").read()
    head, pos = {}, 0
    while True:
        end = raw.index(b"\n", pos)
        line = raw[pos:end].decode()
        pos = end + 1
        if line.startswith("#"):
            continue
        key, _, val = line.partition(" ")
        head[key] = val
        if key == "DATA":
            return head, raw[pos:]


def test_pcd_rgb_binary_and_ascii(kfx, tmp_path):
    p, c = _points()
    for binary in (1, 0):
        path = str(tmp_path / f"cloud{binary}.pcd")
        kfx.check(kfx.lib.kfx_save_pcd(path.encode(), _ptr(p), _ptr(c), len(p), binary))
        head, body = _read_pcd(path)
        assert head["VERSION"] == "0.7" and head["FIELDS"] == "x y z rgb" and head["SIZE"] == "4 4 4 4"
        assert head["TYPE"] == "F F F F" and head["WIDTH"] == "3" and head["POINTS"] == "3"
        assert head["HEIGHT"] == "1" and head["DATA"] == ("binary" if binary else "ascii")
        if binary:
            rec = np.frombuffer(body, np.float32).reshape(3, 4)
            xyz, rgb = rec[:, :3], rec[:, 3].view(np.uint32)
        else:
            rows = [r.split() for r in body.decode().strip().split("\n")]
            xyz = np.array([[float(v) for v in r[:3]] for r in rows], np.float32)
            rgb = np.array([int(r[3]) for r in rows], np.uint32)
        assert np.array_equal(xyz.view(np.uint32), p[:, :3].view(np.uint32))
        # The reference's packing: r <- colour b, b <- colour r.
        want = (c[:, 2].astype(np.uint32) << 16) | (c[:, 1].astype(np.uint32) << 8) | c[:, 0]
        assert np.array_equal(rgb, want)


def test_pcd_intensity_without_colours(kfx, tmp_path):
    p, _ = _points()
    path = str(tmp_path / "xyzi.pcd")
    kfx.check(kfx.lib.kfx_save_pcd(path.encode(), _ptr(p), None, len(p), 1))
    head, body = _read_pcd(path)
    assert head["FIELDS"] == "x y z intensity"
    assert np.array_equal(np.frombuffer(body, np.float32).reshape(3, 4), p)


def test_records_file_round_trip(kfx, tmp_path):
    rng = np.random.default_rng(5)
    r = np.zeros(1000, kfx.SLICE_RECORD_DTYPE)
    for f in ("x", "y", "z"):
        r[f] = rng.integers(-5000, 5000, r.size)
    r["tsdf"] = rng.integers(-32767, 32768, r.size)
    r["weight"] = rng.integers(0, 129, r.size)
    path = str(tmp_path / "slice.rec")
    kfx.check(kfx.lib.kfx_save_records(path.encode(), _ptr(r), r.size))
    raw = open(path, "rb").read()
    assert raw[:8] == b"KFXREC1\0"
    assert int.from_bytes(raw[8:16], "little") == r.size
    assert np.array_equal(np.frombuffer(raw[16:], kfx.SLICE_RECORD_DTYPE), r)
    assert kfx.lib.kfx_save_records(path.encode(), None, 3) != 0
