"""MI355X C ABI vs the CPU oracle on identical inputs (run with -m gpu).

Bar: bit-exact for every stencil, fill, boundary and update (same operation
order, no fma on either side); reductions to 1e-12 relative (double sums in a
different order); min / max / maxErr exact.
"""
import numpy as np
import pytest

from hpcscan_amd.layout import make_layout

pytestmark = pytest.mark.gpu

PRECS = {  # name: (storage dtype, compute fp64)
    "f32": (np.float32, False),
    "f64": (np.float64, True),
    "f32c64": (np.float32, True),
}
SHAPES = {1: [(37,)], 2: [(23, 19)], 3: [(21, 18, 17), (9, 10, 11)]}


def _layouts(oracle, dim, n, order, dtype, **kw):
    fp64 = dtype == np.float64
    p = make_layout(dim, n, order, fp64=fp64, **kw)
    o = oracle.layout(dim, n, order, storage_fp64=fp64, **kw)
    assert p.npoint == o.npoint
    return p, o


def _rand(rng, n, dtype):
    return rng.standard_normal(n).astype(dtype)


def _same(a, b):
    """bitwise equality up to the sign of zero"""
    assert a.dtype == b.dtype and a.shape == b.shape
    eq = (a == b) | (np.isnan(a) & np.isnan(b))
    if not eq.all():
        i = np.flatnonzero(~eq)[:5]
        raise AssertionError(f"{(~eq).sum()} values differ, e.g. at {i}: {a[i]} vs {b[i]}")


@pytest.fixture(params=["ref", "auto"])
def family(request, capi):
    capi.set_kernel(capi.HSX_KERNEL_REF if request.param == "ref" else capi.HSX_KERNEL_AUTO)
    yield request.param
    capi.set_kernel(capi.HSX_KERNEL_AUTO)


def _cases():
    out = []
    for dim, shapes in SHAPES.items():
        for n in shapes:
            for order in (2, 4, 6, 8, 10, 12, 14, 16):
                if min(n[:dim]) < order // 2 + 2:
                    continue
                for prec in PRECS:
                    out.append((dim, n, order, prec))
    return out


@pytest.mark.parametrize("dim,n,order,prec", _cases())
def test_stencils_bit_exact(oracle, capi, family, dim, n, order, prec):
    dtype, c64 = PRECS[prec]
    p, o = _layouts(oracle, dim, n, order, dtype)
    print(f"case dim={dim} n={n} order={order} prec={prec} family={family}", flush=True)
    rng = np.random.default_rng(1000 * dim + order)
    g = capi.hsx_grid(p, compute_fp64=c64)
    u = _rand(rng, p.npoint, dtype)
    coef = (0.1 * _rand(rng, p.npoint, dtype)).astype(dtype)
    prn0 = _rand(rng, p.npoint, dtype)
    w0 = _rand(rng, p.npoint, dtype)
    iv = oracle.inv2(p.d, dtype, compute64=c64)
    du, dw = capi.DeviceArray.from_numpy(u), capi.DeviceArray.from_numpy(w0)
    dc, dp = capi.DeviceArray.from_numpy(coef), capi.DeviceArray.from_numpy(prn0)

    for axis in range(dim):
        ref = w0.copy()
        oracle.fd_d2(o, axis, u, ref, iv, compute64=c64)
        dw.upload(w0)
        capi.fd_d2(g, axis, du, dw)
        _same(dw.download(dtype), ref)
        ref = w0.copy()
        oracle.fd_d1(o, axis, u, ref, oracle.inv1(p.d, dtype, compute64=c64), compute64=c64)
        dw.upload(w0)
        capi.fd_d1(g, axis, du, dw)
        _same(dw.download(dtype), ref)

    ref = w0.copy()
    oracle.laplacian(o, u, ref, iv, compute64=c64)
    dw.upload(w0)
    capi.laplacian(g, du, dw)
    _same(dw.download(dtype), ref)

    ref = prn0.copy()
    oracle.pressure_fd(o, ref, u, coef, iv, compute64=c64)
    capi.ac2_step(g, dp, du, dc)
    _same(dp.download(dtype), ref)

    # SplitComp update with w0 as the Laplacian grid
    ref = prn0.copy()
    oracle.update_pressure(o, ref, u, coef, w0, compute64=c64)
    dp.upload(prn0)
    dw.upload(w0)
    capi.update_pressure(g, dp, du, dc, dw)
    _same(dp.download(dtype), ref)


@pytest.mark.parametrize("prec", list(PRECS))
@pytest.mark.parametrize("dim,n,order,mul", [
    (3, (21, 18, 17), 2, 4), (3, (21, 18, 17), 6, 8), (3, (9, 10, 11), 8, 4), (3, (40, 6, 5), 16, 4),
    (2, (23, 19), 4, 4), (1, (37,), 10, 8), (3, (21, 18, 17), 8, 0)])
def test_update_pressure_wide_bit_exact(oracle, capi, dim, n, order, mul, prec):
    """updatePressure on grids whose n1 is a multiple of the 16-byte vector
    (n1MulPad), where the 16-byte-lane kernel runs and the edge groups of each
    row hold points outside the box; mul 0: the reference's own n1 padding"""
    dtype, c64 = PRECS[prec]
    mp = (mul, -1, -1) if mul else (-1, -1, -1)
    p, o = _layouts(oracle, dim, n, order, dtype, mul_pad=mp)
    print(f"case dim={dim} n={n} order={order} prec={prec} n1={p.n1} inner={p.inner()}", flush=True)
    rng = np.random.default_rng(7 * order + dim)
    g = capi.hsx_grid(p, compute_fp64=c64)
    u, coef, prn0, lap = (_rand(rng, p.npoint, dtype) for _ in range(4))
    ref = prn0.copy()
    oracle.update_pressure(o, ref, u, coef, lap, compute64=c64)
    dp = capi.DeviceArray.from_numpy(prn0)
    du, dc, dl = (capi.DeviceArray.from_numpy(x) for x in (u, coef, lap))
    capi.update_pressure(g, dp, du, dc, dl)
    _same(dp.download(dtype), ref)


@pytest.mark.parametrize("dim,n", [(1, (30,)), (2, (17, 20)), (3, (12, 11, 13))])
@pytest.mark.parametrize("order", [2, 8, 16])
@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_free_surface_bit_exact(oracle, capi, dim, n, order, prec):
    dtype, _ = PRECS[prec]
    p, o = _layouts(oracle, dim, n, order, dtype)
    rng = np.random.default_rng(order)
    u = _rand(rng, p.npoint, dtype)
    ref = u.copy()
    oracle.free_surface(o, ref)
    d = capi.DeviceArray.from_numpy(u)
    capi.free_surface(capi.hsx_grid(p), d)
    _same(d.download(dtype), ref)


@pytest.mark.parametrize("n,order,prec", [((21, 11, 13), 2, "f32"), ((21, 12, 10), 6, "f32"), ((22, 11, 12), 6, "f64"),
                                          ((22, 9, 11), 2, "f64"), ((20, 9, 12), 4, "f32")])
@pytest.mark.parametrize("mask", [0x3F, 0x3C, 0x0C, 0x30, 0x33])
def test_free_surface_unaligned_n1_ranges(oracle, capi, n, order, prec, mask):
    """n1 a multiple of the 16-byte vector but the inner n1 range not aligned to it:
    the n2 / n3 faces run in 16-byte groups whose edge groups hold points
    outside the face (left alone), and the n1 faces (where R allows) take the
    row-end kernel"""
    dtype, _ = PRECS[prec]
    p, o = _layouts(oracle, 3, n, order, dtype)
    vw = 16 // np.dtype(dtype).itemsize
    print(f"free surface n={n} order={order} prec={prec} mask={mask:#x} n1={p.n1} inner={p.inner()[:2]}", flush=True)
    assert p.n1 % vw == 0
    u = _rand(np.random.default_rng(mask + order), p.npoint, dtype)
    ref = u.copy()
    oracle.free_surface(o, ref, face_mask=mask)
    d = capi.DeviceArray.from_numpy(u)
    capi.free_surface(capi.hsx_grid(p), d, face_mask=mask)
    _same(d.download(dtype), ref)


@pytest.mark.parametrize("n,order", [((7, 9, 6), 8), ((5, 12, 11), 8), ((9, 10, 11), 16), ((4, 4, 4), 2)])
def test_free_surface_small_grids(oracle, capi, n, order):
    """axes with fewer than R+2 inner points: the six phases run one after the other"""
    p, o = _layouts(oracle, 3, n, order, np.float32)
    u = _rand(np.random.default_rng(11), p.npoint, np.float32)
    ref = u.copy()
    oracle.free_surface(o, ref)
    d = capi.DeviceArray.from_numpy(u)
    capi.free_surface(capi.hsx_grid(p), d)
    _same(d.download(np.float32), ref)


@pytest.mark.parametrize("mask", [0x01, 0x02, 0x0C, 0x30, 0x15, 0x2A, 0x3E, 0x1F, 0x03, 0x00])
@pytest.mark.parametrize("prec", ["f32", "f64"])
@pytest.mark.parametrize("n", [(14, 12, 13), (16, 12, 13)])
def test_free_surface_face_masks(oracle, capi, mask, prec, n):
    """faces with a neighbour left out (multi-rank decomposition); n1 = 24 in
    FP32 takes the 16-byte n1-face kernel, n1 = 22 the per-point one"""
    dtype, _ = PRECS[prec]
    p, o = _layouts(oracle, 3, n, 8, dtype)
    u = _rand(np.random.default_rng(mask), p.npoint, dtype)
    ref = u.copy()
    oracle.free_surface(o, ref, face_mask=mask)
    d = capi.DeviceArray.from_numpy(u)
    capi.free_surface(capi.hsx_grid(p), d, face_mask=mask)
    _same(d.download(dtype), ref)


def test_free_surface_with_layers_and_padding(oracle, capi):
    for kw in (dict(nlayer=2), dict(add_pad=(3, 2, 1)), dict(offset=(1, 2, 0), mul_pad=(5, 5, 5)),
               dict(auto_pad=True)):
        p, o = _layouts(oracle, 3, (14, 12, 13), 6, np.float32, **kw)
        u = _rand(np.random.default_rng(3), p.npoint, np.float32)
        ref = u.copy()
        oracle.free_surface(o, ref)
        d = capi.DeviceArray.from_numpy(u)
        capi.free_surface(capi.hsx_grid(p), d)
        _same(d.download(np.float32), ref)


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_fills_bit_exact(oracle, capi, prec):
    dtype, _ = PRECS[prec]
    p, o = _layouts(oracle, 3, (20, 17, 15), 8, dtype)
    g = capi.hsx_grid(p)
    base = np.zeros(p.npoint, dtype)
    d = capi.DeviceArray.from_numpy(base)
    for ptype in (oracle.INNER_POINTS, oracle.ALL_POINTS, oracle.I2HALO1, oracle.MIDDLE_POINT):
        r = oracle.index_range(o, ptype)
        ref = base.copy()
        oracle.fill_const(o, ref, ptype, 0.7)
        d.upload(base)
        capi.fill_const(g, d, r, dtype(0.7))
        _same(d.download(dtype), ref)
    # separable sine fill (Grid::fill with FUNC_SINE on every axis)
    orig = [oracle.unit_grid(20 if a == 0 else 17 if a == 1 else 15, o.ax[a].inner_start,
                             fp64=dtype == np.float64)[1] for a in range(3)]
    ref = base.copy()
    tabs = oracle.fill_sep(o, ref, oracle.ALL_POINTS, (1, 2, 1), (2.1, 0.7, 3.3), -1.3, orig, p.d, 3)
    d.upload(base)
    capi.fill_separable(g, d, oracle.index_range(o, oracle.ALL_POINTS), *tabs, -1.3)
    _same(d.download(dtype), ref)


@pytest.mark.parametrize("prec", ["f32", "f64"])
@pytest.mark.parametrize("n", [(3, 5, 4), (7, 9, 6), (1100, 5, 4)])
@pytest.mark.parametrize("shift", [0, 1])
def test_fills_vector_and_point_paths(oracle, capi, prec, n, shift):
    """16-byte stores with point-by-point row heads and tails when the grid is
    16-byte aligned (shift 0), point by point otherwise (shift 1); the element
    before a shifted grid stays untouched"""
    dtype, _ = PRECS[prec]
    it = np.dtype(dtype).itemsize
    p, o = _layouts(oracle, 3, n, 4, dtype)
    g = capi.hsx_grid(p)
    base = np.zeros(p.npoint + shift, dtype)
    base[:shift] = dtype(5.5)
    d = capi.DeviceArray.from_numpy(base)
    pu = d.ptr.value + shift * it

    def got():
        out = d.download(dtype)
        assert (out[:shift] == dtype(5.5)).all()
        return out[shift:]

    for ptype in (oracle.INNER_POINTS, oracle.ALL_POINTS, oracle.I2HALO1, oracle.MIDDLE_POINT):
        ref = base[shift:].copy()
        oracle.fill_const(o, ref, ptype, 0.7)
        d.upload(base)
        capi.fill_const(g, pu, oracle.index_range(o, ptype), dtype(0.7))
        _same(got(), ref)
    orig = [oracle.unit_grid(n[a], o.ax[a].inner_start, fp64=dtype == np.float64)[1] for a in range(3)]
    for ptype in (oracle.INNER_POINTS, oracle.ALL_POINTS):
        ref = base[shift:].copy()
        tabs = oracle.fill_sep(o, ref, ptype, (1, 2, 1), (2.1, 0.7, 3.3), -1.3, orig, p.d, 3)
        d.upload(base)
        capi.fill_separable(g, pu, oracle.index_range(o, ptype), *tabs, -1.3)
        _same(got(), ref)


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_reductions(oracle, capi, prec):
    dtype, _ = PRECS[prec]
    p, o = _layouts(oracle, 3, (40, 33, 21), 8, dtype)
    rng = np.random.default_rng(11)
    u = _rand(rng, p.npoint, dtype)
    v = _rand(rng, p.npoint, dtype)
    v[::7] = dtype(1e-7)  # exercise the |ref| < MAX_ERR_FLOAT branch of maxErr
    g = capi.hsx_grid(p)
    du, dv = capi.DeviceArray.from_numpy(u), capi.DeviceArray.from_numpy(v)
    maxerrf = 1e-11 if dtype == np.float64 else 1e-5
    for ptype in (oracle.INNER_POINTS, oracle.ALL_POINTS):
        r = oracle.index_range(o, ptype)
        a = capi.sum_abs(g, du, r)
        assert abs(a / oracle.reduce("sum_abs", o, u, ptype=ptype) - 1) < 1e-12
        a = capi.sum_abs_diff(g, du, dv, r)
        assert abs(a / oracle.reduce("sum_abs_diff", o, u, v, ptype=ptype) - 1) < 1e-12
        s1, s2 = capi.l1_sums(g, du, dv, r)
        assert abs(s1 / s2 / oracle.reduce("l1_err", o, u, v, ptype=ptype) - 1) < 1e-12
        assert capi.max_err(g, du, dv, r, maxerrf) == oracle.reduce("max_err", o, u, v, ptype=ptype)
        mn, mx = capi.min_max(g, du, r)
        assert mn == oracle.reduce("min", o, u, ptype=ptype)
        assert mx == oracle.reduce("max", o, u, ptype=ptype)


@pytest.mark.parametrize("prec", ["f32", "f64"])
@pytest.mark.parametrize("n", [(3, 5, 4), (7, 9, 6), (1100, 5, 4)])
@pytest.mark.parametrize("shift", [0, 1])
def test_reductions_vector_and_point_paths(oracle, capi, prec, n, shift):
    """16-byte loads with a point-by-point head and tail on each row when both
    buffers are 16-byte aligned (shift 0), point by point otherwise (shift 1):
    rows shorter than one vector group, rows of a few groups, and rows of
    several 4 x 64-group passes with a partial last pass"""
    dtype, _ = PRECS[prec]
    it = np.dtype(dtype).itemsize
    p, o = _layouts(oracle, 3, n, 4, dtype)
    rng = np.random.default_rng(23)
    u = _rand(rng, p.npoint, dtype)
    v = _rand(rng, p.npoint, dtype)
    v[::5] = dtype(1e-7)
    g = capi.hsx_grid(p)
    pad = np.zeros(shift, dtype)
    du = capi.DeviceArray.from_numpy(np.concatenate([pad, u]))
    dv = capi.DeviceArray.from_numpy(np.concatenate([pad, v]))
    pu, pv = du.ptr.value + shift * it, dv.ptr.value + shift * it
    maxerrf = 1e-11 if dtype == np.float64 else 1e-5
    for ptype in (oracle.INNER_POINTS, oracle.ALL_POINTS):
        r = oracle.index_range(o, ptype)
        assert abs(capi.sum_abs(g, pu, r) / oracle.reduce("sum_abs", o, u, ptype=ptype) - 1) < 1e-12
        assert abs(capi.sum_abs_diff(g, pu, pv, r) / oracle.reduce("sum_abs_diff", o, u, v, ptype=ptype) - 1) < 1e-12
        s1, s2 = capi.l1_sums(g, pu, pv, r)
        assert abs(s1 / s2 / oracle.reduce("l1_err", o, u, v, ptype=ptype) - 1) < 1e-12
        assert capi.max_err(g, pu, pv, r, maxerrf) == oracle.reduce("max_err", o, u, v, ptype=ptype)
        assert capi.min_max(g, pu, r) == (oracle.reduce("min", o, u, ptype=ptype), oracle.reduce("max", o, u, ptype=ptype))


@pytest.mark.parametrize("fp64", [False, True])
@pytest.mark.parametrize("n", [(16, 15, 14), (13, 7, 5)])
def test_array_ops(capi, fp64, n):
    """vector path (16-byte aligned buffers, element tail when npoint is not a
    multiple of the vector width)"""
    p = make_layout(3, n, 4, fp64=fp64)
    g = capi.hsx_grid(p)
    dt = np.float64 if fp64 else np.float32
    rng = np.random.default_rng(5)
    u, v, w = (_rand(rng, p.npoint, dt) for _ in range(3))
    du, dv, dw = (capi.DeviceArray.from_numpy(x) for x in (u, v, w))
    capi.array_op(g, capi.HSX_ARRAY_ADD, dw, du, dv)
    _same(dw.download(dt), u + v)
    capi.array_op(g, capi.HSX_ARRAY_MULTIPLY, dw, du, dv)
    _same(dw.download(dt), u * v)
    capi.array_op(g, capi.HSX_ARRAY_ADDTO, dw, du)
    _same(dw.download(dt), u * v + u)
    capi.array_op(g, capi.HSX_ARRAY_COPY, dw, du)
    _same(dw.download(dt), u)
    capi.array_op(g, capi.HSX_ARRAY_FILL, dw, val=2.5)
    _same(dw.download(dt), np.full_like(u, 2.5))


def test_array_ops_unaligned(capi):
    """buffers offset by one element take the element path"""
    p = make_layout(3, (16, 15, 14), 4)
    g = capi.hsx_grid(p)
    rng = np.random.default_rng(6)
    u, v = (_rand(rng, p.npoint + 1, np.float32) for _ in range(2))
    w = np.zeros(p.npoint + 1, np.float32)
    du, dv, dw = (capi.DeviceArray.from_numpy(x) for x in (u, v, w))
    off = lambda d: d.ptr.value + 4  # noqa: E731
    capi.array_op(g, capi.HSX_ARRAY_ADD, off(dw), off(du), off(dv))
    exp = w.copy()
    exp[1:] = u[1:] + v[1:]
    _same(dw.download(np.float32), exp)


@pytest.mark.parametrize("dim,n,order", [(3, (30, 28, 26), 8), (3, (20, 20, 20), 16), (2, (60, 50), 4),
                                         (1, (90,), 12), (3, (64, 40, 33), 2)])
@pytest.mark.parametrize("split", [False, True])
def test_propagation_bit_exact(oracle, capi, family, dim, n, order, split):
    """30 leapfrog steps (free surface + fused step, or Laplacian + update) on the
    device give the oracle's wavefield bit for bit."""
    print(f"propagation dim={dim} n={n} order={order} split={split} family={family}", flush=True)
    p, o = _layouts(oracle, dim, n, order, np.float32)
    rng = np.random.default_rng(order + dim)
    prc = _rand(rng, p.npoint, np.float32)
    prn = _rand(rng, p.npoint, np.float32)
    coef = np.zeros(p.npoint, np.float32)
    oracle.fill_const(o, coef, oracle.INNER_POINTS, 0.05)
    lap = np.zeros(p.npoint, np.float32)
    iv = oracle.inv2(p.d, np.float32)
    g = capi.hsx_grid(p)
    dprc, dprn, dcoef, dlap = (capi.DeviceArray.from_numpy(x) for x in (prc, prn, coef, lap))
    a, b = prn.copy(), prc.copy()
    for it in range(30):
        oracle.free_surface(o, b)
        capi.free_surface(g, dprc)
        if split:
            oracle.laplacian(o, b, lap, iv)
            oracle.update_pressure(o, a, b, coef, lap)
            capi.laplacian(g, dprc, dlap)
            capi.update_pressure(g, dprn, dprc, dcoef, dlap)
        else:
            oracle.pressure_fd(o, a, b, coef, iv)
            capi.ac2_step(g, dprn, dprc, dcoef)
        a, b = b, a
        dprn, dprc = dprc, dprn
    _same(dprc.download(np.float32), b)
    _same(dprn.download(np.float32), a)


def test_invalid_arguments_are_rejected(capi):
    """Host-side validation: no kernel is launched for a bad descriptor."""
    p = make_layout(3, (10, 10, 10), 8)
    g = capi.hsx_grid(p)
    d = capi.DeviceArray(p.npoint * 4)
    bad = capi.hsx_grid(p)
    bad.fd_order = 9
    with pytest.raises(capi.HsxError):
        capi.ac2_step(bad, d, d, d)
    bad = capi.hsx_grid(p)
    bad.inner[1] = p.n1 - 1  # inner + halo would read past the row
    with pytest.raises(capi.HsxError):
        capi.laplacian(bad, d, d)
    with pytest.raises(capi.HsxError):
        capi.ac2_step(g, d, d, d)  # prn == prc
    with pytest.raises(capi.HsxError):
        capi.fd_d2(g, 3, d, d)
    with pytest.raises(capi.HsxError):
        capi.sum_abs(g, d, [0, p.n1, 0, 0, 0, 0])


def _n1_aligned(target, order, dtype, vw):
    """smallest inner n1 >= target whose padded row length is a multiple of vw"""
    for n in range(target, target + 16):
        if make_layout(3, (n, 8, 8), order, fp64=dtype == np.float64).n1 % vw == 0:
            return n
    raise AssertionError("no aligned size")


TILED_CASES = [
    # (order, prec, n1 target, n2, n3, lanes, chunk, rows, flags); flags 1 cached, 2 nontemporal, 4 no ring
    (8, "f32", 72, 33, 29, 16, 5, 0, 0),      # several tiles in n1 and n2, chunk < 2R+1
    (8, "f32", 72, 33, 29, 16, 9, 0, 2),      # chunk == 2R+1, nontemporal streams
    (8, "f32", 130, 17, 40, 32, 13, 8, 1),
    (8, "f32", 130, 37, 40, 32, 13, 16, 2),   # 16 rows with the halo ring, partial last tile row
    (8, "f32", 130, 37, 40, 32, 13, 16, 6),   # 16 rows without the ring
    (8, "f32", 200, 24, 31, 64, 0, 8, 0),     # one wide tile, default chunk, 8 rows with the ring
    (8, "f32", 200, 37, 31, 64, 0, 8, 4),     # 8 rows without the ring
    (8, "f32", 200, 37, 31, 64, 11, 8, 1),    # 8 rows, ring, cached streams, chunk not a multiple of 2R+1
    (2, "f32", 140, 13, 22, 64, 5, 8, 0),     # R = 1 ring at 8 rows
    (16, "f32", 140, 19, 20, 32, 0, 8, 0),    # R = 8 ring at 8 rows: two n2 halo rows per thread
    (8, "f32", 200, 40, 31, 64, 7, 16, 1),
    (8, "f32", 200, 40, 3, 64, 0, 16, 0),     # fewer planes than R: ring filled past the chunk
    (2, "f32", 69, 20, 22, 16, 7, 0, 0),      # odd halo: n1 offset, tile origin below the inner start
    (2, "f32", 140, 20, 22, 32, 5, 16, 0),    # R = 1 ring (two slots)
    (6, "f32", 69, 19, 23, 16, 4, 0, 0),
    (6, "f32", 140, 35, 23, 32, 4, 16, 2),
    (16, "f32", 68, 25, 37, 16, 6, 0, 0),     # R = 8: two n2 halo rows per thread at 8 rows
    (16, "f32", 140, 9, 20, 64, 0, 8, 0),
    (16, "f32", 140, 41, 20, 32, 0, 16, 2),   # R = 8 at 16 rows with the ring
    (16, "f32", 270, 41, 20, 64, 0, 16, 0),   # R = 8, 64 lanes: ring too large, 16 rows without it
    (16, "f32", 200, 41, 20, 0, 0, 0, 0),     # default shape at R = 8 FP32, 64-lane width: 32 x 16 ring
    (14, "f32", 196, 37, 23, 0, 0, 0, 0),     # the same at R = 7 (odd: n1 offset)
    (16, "f32", 500, 19, 37, 0, 0, 0, 0),     # four 128-point tiles per row, last one partial
    (8, "f32", 300, 20, 20, 0, 0, 0, 0),      # default at 300 points per row: 32-lane tiles (three of 128)
    (8, "f64", 250, 21, 19, 0, 0, 0, 0),      # default FP64 at order 8: 64 x 8 with the ring, nontemporal
    (8, "f32c64", 250, 21, 19, 0, 0, 0, 0),   # FP32 storage, FP64 compute: 64 x 8 without the ring
    (4, "f64", 70, 21, 27, 16, 5, 0, 0),      # VW = 2
    (4, "f64", 150, 37, 27, 64, 9, 16, 2),
    (4, "f64", 150, 37, 27, 64, 9, 16, 4),
    (4, "f64", 150, 19, 27, 64, 0, 8, 0),     # FP64 ring at 8 rows
    (10, "f64", 150, 23, 25, 32, 7, 8, 2),
    (12, "f64", 66, 30, 26, 32, 11, 0, 1),
    (8, "f32c64", 72, 18, 21, 16, 5, 0, 0),   # FP32 storage, FP64 compute
    (8, "f32c64", 140, 35, 21, 64, 5, 16, 2),
    (8, "f32", 600, 20, 30, 128, 0, 8, 0),     # 128 x 8 ring: two tiles of 512 points, the second partial
    (8, "f32", 300, 21, 19, 128, 0, 0, 0),     # 128 lanes, rows left to the library (8)
    (8, "f32", 560, 21, 40, 128, 7, 8, 0),     # 128 x 8 ring in 7-plane chunks
    (6, "f32", 520, 19, 23, 128, 0, 8, 0),     # R = 3 (n1 offset)
    (2, "f32", 700, 17, 20, 128, 0, 8, 2),     # R = 1, nontemporal streams
    (4, "f32", 520, 20, 20, 128, 0, 8, 0),     # R = 2: the instance would spill, library choice instead
    (16, "f32", 260, 40, 36, 128, 0, 8, 0),    # R = 8: ring too large, library choice
    (8, "f64", 300, 20, 22, 128, 0, 8, 0),     # FP64: library choice
    (8, "f32", 560, 21, 40, 128, 0, 8, 4),     # 128 lanes without the ring: library choice
    (8, "f32", 480, 21, 19, 0, 0, 0, 0),       # library choice: 128 x 8 (one 512-point tile per row)
    (8, "f32", 960, 13, 17, 0, 0, 0, 0),       # library choice: 128 x 8 (two tiles, 94 % of the width)
    (8, "f32", 800, 13, 17, 0, 0, 0, 0),       # library choice: 64 x 16 (128 lanes would fill 78 %)
]


@pytest.mark.parametrize("order,prec,n1,n2,n3,lanes,chunk,rows,flags", TILED_CASES)
def test_tiled_shapes_bit_exact(oracle, capi, order, prec, n1, n2, n3, lanes, chunk, rows, flags):
    """The LDS-tiled kernels, forced, on partial tiles and chunk boundaries."""
    dtype, c64 = PRECS[prec]
    vw = 16 // np.dtype(dtype).itemsize
    n = (_n1_aligned(n1, order, dtype, vw), n2, n3)
    streams = flags
    print(f"tiled order={order} prec={prec} n={n} lanes={lanes} chunk={chunk} rows={rows} "
          f"flags={flags}", flush=True)
    p, o = _layouts(oracle, 3, n, order, dtype)
    assert p.n1 % vw == 0
    rng = np.random.default_rng(order * 7 + n3)
    g = capi.hsx_grid(p, compute_fp64=c64)
    u, coef, prn0, w0 = (_rand(rng, p.npoint, dtype) for _ in range(4))
    iv = oracle.inv2(p.d, dtype, compute64=c64)
    du, dc, dp, dw = (capi.DeviceArray.from_numpy(x) for x in (u, coef, prn0, w0))
    capi.set_kernel(capi.HSX_KERNEL_TILED)
    capi.set_tile(lanes, chunk, rows, streams)
    try:
        iv1 = oracle.inv1(p.d, dtype, compute64=c64)
        for axis in range(3):
            ref = w0.copy()
            oracle.fd_d2(o, axis, u, ref, iv, compute64=c64)
            dw.upload(w0)
            capi.fd_d2(g, axis, du, dw)
            _same(dw.download(dtype), ref)
            ref = w0.copy()
            oracle.fd_d1(o, axis, u, ref, iv1, compute64=c64)
            dw.upload(w0)
            capi.fd_d1(g, axis, du, dw)
            _same(dw.download(dtype), ref)
        ref = w0.copy()
        oracle.laplacian(o, u, ref, iv, compute64=c64)
        dw.upload(w0)
        capi.laplacian(g, du, dw)
        _same(dw.download(dtype), ref)
        ref = prn0.copy()
        oracle.pressure_fd(o, ref, u, coef, iv, compute64=c64)
        capi.ac2_step(g, dp, du, dc)
        _same(dp.download(dtype), ref)
    finally:
        capi.set_tile(0, 0)
        capi.set_kernel(capi.HSX_KERNEL_AUTO)


@pytest.mark.parametrize("args", [(8, 0, 0, 0), (64, -1, 0, 0), (64, 0, 12, 0), (64, 0, 0, 3), (64, 0, 0, 8),
                                  (128, 0, 16, 0), (256, 0, 8, 0)])
def test_set_tile_rejects_bad_shapes(capi, args):
    lanes, chunk, rows, flags = args
    with pytest.raises(capi.HsxError):
        capi.set_tile(lanes, chunk, rows, flags)


def test_tiled_rejects_misaligned_rows(capi):
    """Forced tiled kernels refuse rows that are not a multiple of 16 bytes."""
    p = make_layout(3, (21, 18, 17), 8)   # n1 = 29
    g = capi.hsx_grid(p)
    d = capi.DeviceArray(p.npoint * 4)
    capi.set_kernel(capi.HSX_KERNEL_TILED)
    try:
        with pytest.raises(capi.HsxError):
            capi.laplacian(g, d, capi.DeviceArray(p.npoint * 4))
    finally:
        capi.set_kernel(capi.HSX_KERNEL_AUTO)


FOLD_CASES = [
    # (order, prec, n1 target, n2, n3, face mask, lanes, chunk, rows, flags)
    (8, "f32", 72, 40, 36, 0x3F, 0, 0, 0, 0),        # default tile (16 rows, ring)
    (8, "f32", 72, 40, 36, 0x16, 16, 5, 8, 0),       # some faces only, several chunks per n3 column
    (8, "f32", 200, 37, 31, 0x3F, 64, 7, 16, 4),     # 16 rows without the ring
    (8, "f32", 130, 19, 30, 0x29, 32, 0, 8, 1),      # 8 rows with the ring, cached streams
    (16, "f32", 140, 30, 24, 0x3F, 32, 0, 16, 0),    # R = 8
    (16, "f32", 68, 25, 19, 0x3F, 16, 6, 0, 0),      # R = 8, 8 rows: halo rows of the second tile row are inner
    (2, "f32", 69, 20, 22, 0x3F, 0, 0, 0, 0),        # R = 1: the first lane holds halo and inner columns
    (6, "f32", 69, 19, 23, 0x3F, 16, 4, 0, 0),       # R = 3, n1 offset
    (8, "f32", 72, 6, 6, 0x3F, 0, 0, 0, 0),          # N2 = N3 = R + 2: both mirrors of a plane meet
    (4, "f64", 70, 21, 27, 0x3F, 0, 0, 0, 0),
    (12, "f64", 66, 30, 26, 0x1B, 32, 11, 0, 1),
    (8, "f32c64", 72, 18, 21, 0x3F, 0, 0, 0, 0),
]


@pytest.mark.parametrize("order,prec,n1,n2,n3,mask,lanes,chunk,rows,flags", FOLD_CASES)
def test_folded_free_surface_bit_exact(oracle, capi, order, prec, n1, n2, n3, mask, lanes, chunk, rows, flags):
    """hsx_ac2_step_fs over 6 steps: every new wavefield equals the oracle's free
    surface + fused step (inner points, bit for bit), and once the pending free
    surface is applied, whole grids (halos included) equal the oracle's."""
    dtype, c64 = PRECS[prec]
    vw = 16 // np.dtype(dtype).itemsize
    n = (_n1_aligned(n1, order, dtype, vw), n2, n3)
    print(f"fold order={order} prec={prec} n={n} mask={mask:#x} lanes={lanes} chunk={chunk} rows={rows} "
          f"flags={flags}", flush=True)
    p, o = _layouts(oracle, 3, n, order, dtype)
    rng = np.random.default_rng(order * 11 + n3)
    prc, prn = _rand(rng, p.npoint, dtype), _rand(rng, p.npoint, dtype)
    coef = (0.05 * (1 + 0.5 * rng.random(p.npoint))).astype(dtype)
    iv = oracle.inv2(p.d, dtype, compute64=c64)
    g = capi.hsx_grid(p, compute_fp64=c64)
    dprc, dprn, dcoef = (capi.DeviceArray.from_numpy(x) for x in (prc, prn, coef))
    a, b = prn.copy(), prc.copy()
    capi.set_tile(lanes, chunk, rows, flags)
    try:
        capi.free_surface(g, dprc, face_mask=mask)       # the first free surface is applied, later ones folded
        for it in range(6):
            oracle.free_surface(o, b, compute64=c64, face_mask=mask)
            oracle.pressure_fd(o, a, b, coef, iv, compute64=c64)
            fold = capi.HSX_FOLD_PRC | capi.HSX_FOLD_HALO | (capi.HSX_FOLD_PRN if it >= 1 else 0)
            capi.ac2_step_fs(g, dprn, dprc, dcoef, fold, face_mask=mask)
            got = dprn.download(dtype)
            box = p.inner()
            sl = (slice(box[4], box[5] + 1), slice(box[2], box[3] + 1), slice(box[0], box[1] + 1))
            shape = (p.n3, p.n2, p.n1)
            _same(got.reshape(shape)[sl], a.reshape(shape)[sl])
            a, b = b, a
            dprn, dprc = dprc, dprn
        # apply the pending free surface of both grids, then compare everything
        oracle.free_surface(o, b, compute64=c64, face_mask=mask)
        capi.free_surface(g, dprc, face_mask=mask)
        capi.free_surface(g, dprn, face_mask=mask)
        assert np.array_equal(dprc.download(dtype).view(np.uint8), b.view(np.uint8))
        assert np.array_equal(dprn.download(dtype).view(np.uint8), a.view(np.uint8))
    finally:
        capi.set_tile(0, 0)


def test_folded_free_surface_refuses_small_axes(capi):
    """Fewer than R+2 inner points on a folded axis: no folded kernel."""
    p = make_layout(3, (72, 40, 5), 8)
    g = capi.hsx_grid(p)
    d = [capi.DeviceArray(p.npoint * 4) for _ in range(3)]
    with pytest.raises(capi.HsxError) as e:
        capi.ac2_step_fs(g, d[0], d[1], d[2], capi.HSX_FOLD_PRC)
    assert e.value.rc == capi.HSX_ERR_UNSUPPORTED
    capi.ac2_step_fs(g, d[0], d[1], d[2], capi.HSX_FOLD_PRC, face_mask=0x0F)   # n3 faces not folded


def _step_ms(capi, g, dp, du, dc, reps):
    """median HIP-event time of the fused step (library stream)"""
    ev = [(capi.Event(), capi.Event()) for _ in range(reps)]
    for e0, e1 in ev:
        e0.record()
        capi.ac2_step(g, dp, du, dc)
        e1.record()
    capi.sync()
    return sorted(e0.elapsed_ms(e1) for e0, e1 in ev)[reps // 2]


@pytest.mark.parametrize("n,order,prec", [((400, 400, 400), 8, "f32"), ((320, 300, 280), 16, "f32"),
                                          ((256, 240, 200), 8, "f64")])
def test_autotuned_shape_bit_exact_and_not_slower(oracle, capi, n, order, prec):
    """hsx_autotune picks a tile shape for the layout (its sweep includes the
    library's default shape); the fused step then runs with it, bit for bit as
    the oracle, in no more time than the default shape (the Laplacian keeps
    its own shape and stays exact)"""
    dtype, c64 = PRECS[prec]
    p, o = _layouts(oracle, 3, n, order, dtype)
    g = capi.hsx_grid(p, compute_fp64=c64)
    rng = np.random.default_rng(order + n[0])
    u, coef, prn0 = (_rand(rng, p.npoint, dtype) for _ in range(3))
    du, dc, dp = (capi.DeviceArray.from_numpy(x) for x in (u, coef, prn0))
    capi.autotune_clear()
    try:
        default_ms = _step_ms(capi, g, dp, du, dc, 9)
        shape, tune_ms = capi.autotune(g, reps=5)
        tuned_ms = _step_ms(capi, g, dp, du, dc, 9)
        print(f"n={n} O{order} {prec}: tuned shape {shape} ({tune_ms:.4f} ms in the sweep), "
              f"median {tuned_ms:.4f} ms against {default_ms:.4f} ms for the default", flush=True)
        assert tuned_ms <= default_ms * 1.05 + 0.002
        dp.upload(prn0)
        ref = prn0.copy()
        oracle.pressure_fd(o, ref, u, coef, oracle.inv2(p.d, dtype, compute64=c64), compute64=c64)
        capi.ac2_step(g, dp, du, dc)
        _same(dp.download(dtype), ref)
        w = np.zeros(p.npoint, dtype)
        dw = capi.DeviceArray.from_numpy(w)
        oracle.laplacian(o, u, w, oracle.inv2(p.d, dtype, compute64=c64), compute64=c64)
        capi.laplacian(g, du, dw)
        _same(dw.download(dtype), w)
    finally:
        capi.autotune_clear()


@pytest.fixture
def contracted(capi):
    capi.set_contract(True)
    yield
    capi.set_contract(False)


def _rel_l2(a, b):
    a, b = a.astype(np.float64), b.astype(np.float64)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("order", [4, 8, 12, 16])
@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_contracted_kernels_close_to_oracle(oracle, capi, contracted, order, prec):
    """hsx_set_contract(1): the tiled fused step and Laplacian with fused
    multiply-adds, one application on random fields, against the exact oracle:
    within rounding (relative L2 below 1e-6 in FP32, 1e-14 in FP64) and no
    longer equal (the contraction is active); FD_D2 along one axis stays exact"""
    dtype, c64 = PRECS[prec]
    vw = 16 // np.dtype(dtype).itemsize
    n = (_n1_aligned(72, order, dtype, vw), 40, 36)
    p, o = _layouts(oracle, 3, n, order, dtype)
    g = capi.hsx_grid(p, compute_fp64=c64)
    rng = np.random.default_rng(order * 3 + (1 if c64 else 0))
    u, coef, prn0 = (_rand(rng, p.npoint, dtype) for _ in range(3))
    iv = oracle.inv2(p.d, dtype, compute64=c64)
    du, dc, dp = (capi.DeviceArray.from_numpy(x) for x in (u, coef, prn0))
    inner = np.zeros((p.n3, p.n2, p.n1), bool)
    r = p.inner()
    inner[r[4]:r[5] + 1, r[2]:r[3] + 1, r[0]:r[1] + 1] = True
    inner = inner.ravel()
    tol = 1e-6 if dtype == np.float32 else 1e-14
    ref = prn0.copy()
    oracle.pressure_fd(o, ref, u, coef, iv, compute64=c64)
    capi.ac2_step(g, dp, du, dc)
    got = dp.download(dtype)
    e1 = _rel_l2(got[inner], ref[inner])
    w = np.zeros(p.npoint, dtype)
    dw = capi.DeviceArray.from_numpy(w)
    oracle.laplacian(o, u, w, iv, compute64=c64)
    capi.laplacian(g, du, dw)
    lap = dw.download(dtype)
    e2 = _rel_l2(lap[inner], w[inner])
    print(f"contracted O{order} {prec}: fused step rel L2 {e1:.2e}, Laplacian {e2:.2e}", flush=True)
    assert 0 < e1 < tol and 0 < e2 < tol
    w1 = np.zeros(p.npoint, dtype)
    oracle.fd_d2(o, 2, u, w1, iv, compute64=c64)
    dw.upload(np.zeros(p.npoint, dtype))
    capi.fd_d2(g, 2, du, dw)
    _same(dw.download(dtype), w1)


@pytest.mark.parametrize("order", [12, 16])
def test_contracted_propagation_within_1e5(oracle, capi, contracted, order):
    """the eigen-mode propagation (free surface + contracted fused step, FP32,
    100^3, 100 steps) against the exact oracle's wavefield: relative L2 within
    the 1e-5 bar of the FP32 comparisons (1.2e-6 measured at order 12), and
    the errTestCase within 3 % (it measures the small gap to the eigen mode,
    which the float rounding of the stencil sums moves by about 1 %)"""
    import math

    from hpcscan_amd.layout import unit_origin

    n, modes, nt = (100, 100, 100), (2, 2, 2), 100
    p = make_layout(3, n, order)
    g = capi.hsx_grid(p)
    prm = oracle.propa_params(list(n), order, nt_cfg=nt)
    dt = float(prm["dt"])
    orig = [unit_origin(p, a) for a in range(3)]
    gamma = math.sqrt(sum(m * m for m in modes))
    r_all, r_in = p.all_points(), p.inner()
    pi = oracle.PI_REF
    tabs = [np.array([math.sin((orig[a] + i * p.d[a]) * (pi * modes[a]))
                      for i in range(r_all[2 * a], r_all[2 * a + 1] + 1)]) for a in range(3)]

    def amp(t):
        return math.sin(pi * float(np.float32(t)) * gamma)

    prn, prc, coef, ref = (capi.DeviceArray(p.npoint * 4) for _ in range(4))
    for x in (prn, prc, coef, ref):
        capi.fill_const(g, x, [0, p.n1 - 1, 0, p.n2 - 1, 0, p.n3 - 1], 0.0)
    capi.fill_const(g, coef, r_in, float(np.float32(np.float32(dt) * np.float32(dt))))
    capi.fill_separable(g, prn, r_all, *tabs, amp(-2.0 * dt))
    capi.fill_separable(g, prc, r_all, *tabs, amp(-dt))
    s1 = s2 = np.float32(0)
    for it in range(prm["nt"]):
        capi.free_surface(g, prc)
        capi.ac2_step(g, prn, prc, coef)
        capi.fill_separable(g, ref, r_all, *tabs, amp(it * np.float32(dt)))
        s1 = np.float32(s1 + np.float32(capi.sum_abs_diff(g, prn, ref, r_in)))
        s2 = np.float32(s2 + np.float32(capi.sum_abs(g, ref, r_in)))
        prn, prc = prc, prn
    err_gpu = float(s1 / s2)
    field = prc.download(np.float32)
    err_cpu, prn_cpu, _, _ = oracle.propa_eigen(list(n), order, modes, nt_cfg=nt)
    inner = np.zeros((p.n3, p.n2, p.n1), bool)
    inner[r_in[4]:r_in[5] + 1, r_in[2]:r_in[3] + 1, r_in[0]:r_in[1] + 1] = True
    e = _rel_l2(field[inner.ravel()], prn_cpu[inner.ravel()])
    print(f"contracted O{order} 100^3 x {nt}: wavefield rel L2 {e:.2e}; errTestCase {err_gpu:.6g} vs exact "
          f"{err_cpu:.6g}", flush=True)
    assert e < 1e-5
    assert abs(err_gpu / err_cpu - 1) < 3e-2
