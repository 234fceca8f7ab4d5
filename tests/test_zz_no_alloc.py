"""chs_zdotc_noalloc, the allocation-free variant of chs_zdotc (its partial
sums live in a fixed device array instead of context workspaces), against
chs_zdotc through the C-ABI on two calls in a row, the second with a
smaller input: the same kernels in the same reduction order, so the results
must be bit-identical."""
import ctypes as C

import numpy as np
import pytest

import chs

pytestmark = pytest.mark.gpu


def _dot(fn, ctx, n, x, y):
    out = np.zeros(2)
    rc = fn(ctx.h, n, x.ptr, y.ptr, out.ctypes.data_as(C.POINTER(C.c_double)))
    assert rc == 0, chs.lib().chs_last_error(ctx.h)
    return out


def test_zdotc_noalloc_matches_zdotc():
    L = chs.lib()
    L.chs_zdotc_noalloc.restype = C.c_int
    L.chs_zdotc_noalloc.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p,
                                    C.POINTER(C.c_double)]
    ctx = chs.Context(0)
    rng = np.random.default_rng(7)
    for n in (1_000_003, 4099, 0):
        x = rng.standard_normal(max(n, 1)) + 1j * rng.standard_normal(max(n, 1))
        y = rng.standard_normal(max(n, 1)) + 1j * rng.standard_normal(max(n, 1))
        dx, dy = chs.Gpu.from_host(ctx, x), chs.Gpu.from_host(ctx, y)
        ref = _dot(L.chs_zdotc, ctx, n, dx, dy)
        got = _dot(L.chs_zdotc_noalloc, ctx, n, dx, dy)
        assert np.array_equal(got, ref), (n, got, ref)
        want = np.vdot(x[:n], y[:n]) if n else 0.0
        assert abs(complex(*got) - want) <= 1e-12 * max(1.0, abs(want))
        del dx, dy
    ctx.close()
