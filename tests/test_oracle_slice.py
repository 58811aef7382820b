"""Pins the oracle's slice clean/extract against set arithmetic and direct
enumeration on small volumes.  No GPU needed."""
import numpy as np
import pytest

import slice_inputs as si


@pytest.mark.parametrize("n,origin,mn,mx", [
    ((12, 10, 16), (0, 0, 0), (0, 100, 100), (2, 90, 90)),        # x slab, no wrap
    ((12, 10, 16), (7, 3, 11), (15, -5, 100), (18, -1, 90)),      # x slab of wrapped memory
    ((12, 10, 16), (4, 9, 2), (100, 9, 5), (90, 11, 7)),          # y and z slabs
    ((8, 8, 8), (1, 2, 3), (0, 0, 0), (20, 20, 20)),              # everything
    ((8, 8, 8), (1, 2, 3), (1, 1, 1), (0, 0, 0)),                 # nothing
])
def test_extract_matches_enumeration(kfx, orc, n, origin, mn, mx):
    rb, vol, col = si.random_volume(n, origin, seed=sum(n) + origin[0])
    for c in (col, None):
        pts, cols = orc.extract_slice(vol, c, mn, mx, rb)
        bp, bc = si.brute_force_extract(vol, c, mn, mx, rb)
        assert np.array_equal(pts, bp)
        assert np.array_equal(cols, bc)


def test_clean_is_set_arithmetic(kfx, orc):
    rb, vol, col = si.random_volume((12, 10, 16), (5, 6, 7), seed=3)
    mn, mx = (6, 100, 20), (8, 90, 21)
    v2 = vol.copy()
    orc.clean_slice(v2, mn, mx, rb)
    N = (12, 10, 16)
    o = (5, 6, 7)
    for z in range(N[2]):
        for y in range(N[1]):
            for x in range(N[0]):
                a = (x + o[0]) % N[0] + N[0] * ((y + o[1]) % N[1]) + N[0] * N[1] * ((z + o[2]) % N[2])
                zone = mn[0] <= x + o[0] <= mx[0] or mn[2] <= z + o[2] <= mx[2]
                if zone:
                    assert v2[a]["value"] == 0 and v2[a]["weight"] == 0
                else:
                    assert v2[a] == vol[a]
