"""virs_plane_read_noalloc: the allocation-free layer snapshot, through the C
ABI, against the ticketed snapshot (virs_plane_request + virs_plane_fetch)
of the same state, bit for bit; two calls in a row, the second on a smaller
grid (the staging array keeps the first call's contents); and the size
limit of the static staging array."""

import numpy as np
import pytest

from helpers import gpu_context, tiny_problem
from virs_amd import _lib

pytestmark = pytest.mark.gpu

STATIC_PLANE_CELLS = 2 * 1024 * 1024


def _read_noalloc(f, layer):
    nx, ny, nz = f.dims
    out = np.full((nz, nx), np.nan, np.float32)
    rc = f.L.virs_plane_read_noalloc(f._ctx, int(layer), out.ctypes.data_as(_lib.c_fp))
    return rc, out


@pytest.mark.parametrize("precision", ["fp32", "fp16"])
def test_noalloc_plane_equals_ticketed_snapshot(gpu, precision):
    big = tiny_problem(51, dims=(40, 24, 30), steps=50)
    small = tiny_problem(52, dims=(20, 16, 12), steps=50)
    for p in (big, small):  # the second call on a smaller grid
        print(f"noalloc plane {precision} dims={p['dims']}", flush=True)
        f = gpu_context(p, precision)
        f.step(p["steps"])
        peak = 0.0
        for y in (p["dims"][1] // 2, 1, p["dims"][1] - 2):
            rc, a = _read_noalloc(f, y)
            assert rc == 0, f.L.virs_last_error(f._ctx)
            b = f.plane_fetch(f.plane_request(y))
            assert np.isfinite(a).all()
            np.testing.assert_array_equal(a, b)
            peak = max(peak, float(np.abs(a).max()))
        assert peak > 0  # the middle layer holds the running field
        f.close()


def test_noalloc_plane_limits(gpu):
    from virs_amd.fdtd import FDTD
    p = tiny_problem(53, dims=(20, 16, 12), steps=10)
    f = gpu_context(p, "fp32")
    rc, _ = _read_noalloc(f, 16)
    assert rc == -1 and b"out of range" in f.L.virs_last_error(f._ctx)
    f.close()
    # a plane of 2048 x 1025 cells exceeds the 2^21-cell staging array
    g = FDTD((2048, 3, 1025), 0.05)
    assert 2048 * 1025 > STATIC_PLANE_CELLS
    rc, _ = _read_noalloc(g, 1)
    assert rc == -1 and b"exceeds the static staging" in g.L.virs_last_error(g._ctx)
    g.close()
    # the largest plane that fits
    h = FDTD((2048, 3, 1024), 0.05)
    rc, out = _read_noalloc(h, 1)
    assert rc == 0 and (out == 0).all()
    h.close()
