"""hsx_l1_sums_noalloc (partial sums in the library's static device array,
result in a device buffer, no allocation) against hsx_l1_sums: equal sums on
two calls in a row, the second on a smaller range of a smaller grid."""
import ctypes

import numpy as np
import pytest

from hpcscan_amd.layout import make_layout

pytestmark = pytest.mark.gpu


def _noalloc(capi, g, u, v, r, out):
    rr = (ctypes.c_int64 * 6)(*r)
    capi.check(capi.lib().hsx_l1_sums_noalloc(ctypes.byref(g), capi.dptr(u), capi.dptr(v), rr, capi.dptr(out),
                                              None))
    capi.sync()
    return tuple(out.download(np.float64))


@pytest.mark.parametrize("fp64", [False, True])
def test_l1_sums_noalloc_equals_l1_sums(capi, fp64):
    dt = np.float64 if fp64 else np.float32
    rng = np.random.default_rng(12)
    out = capi.DeviceArray(16)
    cases = [((200, 180, 160), 8, None), ((40, 36, 20), 4, (3, 2, 1))]   # second: smaller grid, sub-range
    for n, order, shrink in cases:
        p = make_layout(3, n, order, fp64=fp64)
        g = capi.hsx_grid(p)
        u = rng.standard_normal(p.npoint).astype(dt)
        v = rng.standard_normal(p.npoint).astype(dt)
        du, dv = capi.DeviceArray.from_numpy(u), capi.DeviceArray.from_numpy(v)
        r = list(p.inner())
        if shrink:
            for a in range(3):
                r[2 * a] += shrink[a]
                r[2 * a + 1] -= shrink[a]
        ref = capi.l1_sums(g, du, dv, r)
        got = _noalloc(capi, g, du, dv, r, out)
        print(f"n={n} order={order} fp64={fp64} range={r}: {got} vs {ref}", flush=True)
        assert got == ref
        # and the host's double sums agree to rounding
        m = np.zeros((p.n3, p.n2, p.n1), bool)
        m[r[4]:r[5] + 1, r[2]:r[3] + 1, r[0]:r[1] + 1] = True
        m = m.ravel()
        s1 = np.abs(u[m].astype(np.float64) - v[m].astype(np.float64)).sum()
        s2 = np.abs(v[m].astype(np.float64)).sum()
        assert got[0] == pytest.approx(s1, rel=1e-12) and got[1] == pytest.approx(s2, rel=1e-12)


def test_l1_sums_noalloc_refuses_bad_arguments(capi):
    p = make_layout(3, (20, 18, 16), 4)
    g = capi.hsx_grid(p)
    u = capi.DeviceArray(p.npoint * 4)
    rr = (ctypes.c_int64 * 6)(*p.inner())
    assert capi.lib().hsx_l1_sums_noalloc(ctypes.byref(g), capi.dptr(u), capi.dptr(u), rr, None, None) == \
        capi.HSX_ERR_INVALID
