"""Small cyclic volumes with sign changes, zero weights and sparse colour for
the slice tests."""
import numpy as np

import kfx
import oracle_ctypes as orc


def random_volume(n, origin, seed, color_fraction=0.02):
    rng = np.random.default_rng(seed)
    rb = kfx.round_buffer(n, 1.0, origin=origin)
    L = rb.buffer_length
    vol = np.zeros(L, dtype=kfx.TSDF_DTYPE)
    # smooth-ish field with zero crossings: sum of a plane and noise
    vol["value"] = np.clip(rng.normal(0, 12000, L), -32767, 32767).astype(np.int16)
    vol["value"][rng.random(L) < 0.05] = 32767          # F == 1 (free space)
    vol["weight"] = rng.integers(1, 129, L).astype(np.int16)
    vol["weight"][rng.random(L) < 0.1] = 0              # unobserved
    col = np.zeros(L, dtype=orc.COLOR_DTYPE)
    m = rng.random(L) < color_fraction
    for f in ("r", "g", "b", "w"):
        col[f][m] = rng.integers(0, 256, m.sum()).astype(np.uint8)
    return rb, vol, col


def brute_force_extract(vol, col, mn, mx, rb):
    """Direct enumeration in (chunk, plane, z, y, x) order from the
    predicate definitions, independent of the scan formulation."""
    N = (rb.voxels_size.x, rb.voxels_size.y, rb.voxels_size.z)
    o = (rb.origin_GRID.x, rb.origin_GRID.y, rb.origin_GRID.z)
    step = N[2] // 8 if N[2] >= 8 and N[2] % 8 == 0 else N[2]

    def addr(x, y, z):
        return (x + o[0]) % N[0] + N[0] * ((y + o[1]) % N[1]) + N[0] * N[1] * ((z + o[2]) % N[2])

    def zone(x, y, z):
        return any(mn[k] <= c + o[k] <= mx[k] for k, c in enumerate((x, y, z)))

    def F(x, y, z):
        v = vol[addr(x, y, z)]
        return np.float32(v["value"]) / np.float32(32767), int(v["weight"])

    def cross(a, b):
        (f, w), (fn, wn) = a, b
        return w != 0 and f != 1 and wn != 0 and fn != 1 and ((f > 0 and fn < 0) or (f < 0 and fn > 0))

    pts, cols = [], []
    for z0 in range(0, N[2], step):
        for p in range(4):
            for z in range(z0, z0 + step):
                for y in range(N[1]):
                    for x in range(N[0]):
                        if not zone(x, y, z):
                            continue
                        if p == 0:
                            ok = x + 1 < N[0] and cross(F(x, y, z), F(x + 1, y, z))
                        elif p == 1:
                            ok = y + 1 < N[1] and cross(F(x, y, z), F(x, y + 1, z))
                        elif p == 2:
                            ok = z + 1 < z0 + step and cross(F(x, y, z), F(x, y, z + 1))
                        else:
                            c = col[addr(x, y, z)] if col is not None else None
                            ok = c is not None and any(int(c[f]) for f in ("r", "g", "b", "w"))
                        if ok:
                            q = (x + (p == 0), y + (p == 1), z + (p == 2))
                            pts.append((*q, 1.0))
                            cc = col[addr(*q)] if col is not None else None
                            cols.append(tuple(int(cc[f]) for f in ("r", "g", "b", "w")) if cc is not None
                                        else (0, 0, 0, 0))
    return np.array(pts, np.float32).reshape(-1, 4), np.array(cols, np.uint8).reshape(-1, 4)
