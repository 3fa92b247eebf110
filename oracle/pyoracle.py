"""Python side of the CPU oracle (test infrastructure only).

Only tests/, ``__graft_entry__.smoke()`` and bench.py's ``cpu_baseline`` leg
may import this module.  It provides

* a ctypes binding of ``_build/liboracle.so`` (the C restatement of the
  reference stencil path, see hpcscan_oracle.h),
* the propagator's dt / nt / snapInc arithmetic in float32, restated from
  src/propagator_Ac2.cpp:181-226 and src/grid.cpp:1924-1957,
* the closed-form discrete eigen-mode recurrence that pins the whole
  propagator + error metric (derived from the mathematics: a sine mode with
  the free-surface odd extension is an eigenvector of the FD Laplacian),
* an exact re-derivation of the centred FD D2 weights (Fornberg's algorithm)
  that pins the rational constants of src/fdm.h:276-507.
"""
from __future__ import annotations

import ctypes
import math
import os
from fractions import Fraction

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle.so")

PI_REF = 3.1415926535897          # src/constant.h:12 (truncated on purpose)
MAX_ERR_FLOAT = {False: np.float32(1e-5), True: np.float32(1e-11)}  # src/constant.h:14-20

# Point_type enum order (src/type_def.h:86-88)
INNER_POINTS, ALL_POINTS, MIDDLE_POINT = 0, 1, 2
I1HALO1, I1HALO2, I2HALO1, I2HALO2, I3HALO1, I3HALO2 = 3, 4, 5, 6, 7, 8
# Func_type enum (src/type_def.h:83)
FUNC_NONE, FUNC_SINE, FUNC_COSINE, FUNC_CONST, FUNC_LINEAR = 0, 1, 2, 3, 4


class OrcAxis(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int64) for n in (
        "offset_start", "offset_end", "halo1_start", "halo1_end", "layer1_start",
        "layer1_end", "inner_start", "inner_end", "layer2_start", "layer2_end",
        "halo2_start", "halo2_end", "pad_start", "pad_end", "n")]


class OrcGrid(ctypes.Structure):
    _fields_ = [("dim", ctypes.c_int), ("fd_order", ctypes.c_int), ("halo", ctypes.c_int),
                ("ax", OrcAxis * 3), ("n1", ctypes.c_int64), ("n2", ctypes.c_int64),
                ("n3", ctypes.c_int64), ("npoint", ctypes.c_int64)]


class OrcLayoutOpts(ctypes.Structure):
    _fields_ = [("dim", ctypes.c_int), ("n_inner", ctypes.c_int64 * 3),
                ("fd_order", ctypes.c_int), ("storage_fp64", ctypes.c_int),
                ("nlayer", ctypes.c_int), ("auto_pad", ctypes.c_int),
                ("add_pad", ctypes.c_int64 * 3), ("mul_pad", ctypes.c_int64 * 3),
                ("offset", ctypes.c_int64 * 3)]


class OrcSched(ctypes.Structure):
    _fields_ = [("mode", ctypes.c_int), ("cb1", ctypes.c_int64), ("cb2", ctypes.c_int64),
                ("cb3", ctypes.c_int64)]


class OrcPropaParams(ctypes.Structure):
    _fields_ = [("dim", ctypes.c_int), ("fd_order", ctypes.c_int), ("nt", ctypes.c_int),
                ("snap_inc", ctypes.c_int), ("split_comp", ctypes.c_int),
                ("dt", ctypes.c_double), ("mode", ctypes.c_int * 3),
                ("d", ctypes.c_double * 3), ("orig", ctypes.c_double * 3)]


_lib = None


def lib():
    """Load liboracle.so (built by ``make -C oracle``)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library not built: {LIB_PATH} (run make -C oracle)")
        L = ctypes.CDLL(LIB_PATH)
        P = ctypes.POINTER
        L.orc_layout.argtypes = [P(OrcLayoutOpts), P(OrcGrid)]
        L.orc_index_range.argtypes = [P(OrcGrid), ctypes.c_int, P(ctypes.c_int64)]
        L.orc_d2_coefs.argtypes = [ctypes.c_int, P(ctypes.c_double)]
        L.orc_d1_coefs.argtypes = [ctypes.c_int, P(ctypes.c_double)]
        L.orc_sum_abs_d2_coef.argtypes = [ctypes.c_int]
        L.orc_sum_abs_d2_coef.restype = ctypes.c_double
        L.orc_cfl.argtypes = [ctypes.c_int, ctypes.c_int]
        L.orc_cfl.restype = ctypes.c_float
        L.orc_inv2_f32.argtypes = [ctypes.c_double]
        L.orc_inv2_f32.restype = ctypes.c_float
        L.orc_inv2_f64.argtypes = [ctypes.c_double]
        L.orc_inv2_f64.restype = ctypes.c_double
        L.orc_fill_table.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                     ctypes.c_double, ctypes.c_int64, ctypes.c_int64,
                                     ctypes.c_double, P(ctypes.c_double)]
        vp = ctypes.c_void_p
        for sfx in ("f32", "f64", "f32c64"):
            getattr(L, f"orc_fill_const_{sfx}").argtypes = [P(OrcGrid), vp, P(ctypes.c_int64),
                                                            ctypes.c_double if sfx == "f64" else ctypes.c_float]
            getattr(L, f"orc_fill_sep_{sfx}").argtypes = [P(OrcGrid), vp, P(ctypes.c_int64), vp, vp,
                                                          vp, ctypes.c_double]
            getattr(L, f"orc_free_surface_{sfx}").argtypes = [P(OrcGrid), vp]
            getattr(L, f"orc_free_surface_mask_{sfx}").argtypes = [P(OrcGrid), vp, ctypes.c_int]
            getattr(L, f"orc_fd_d2_{sfx}").argtypes = [P(OrcGrid), ctypes.c_int, vp, vp, vp, OrcSched]
            getattr(L, f"orc_fd_d1_{sfx}").argtypes = [P(OrcGrid), ctypes.c_int, vp, vp, vp, OrcSched]
            getattr(L, f"orc_laplacian_{sfx}").argtypes = [P(OrcGrid), vp, vp, vp, OrcSched]
            getattr(L, f"orc_update_pressure_{sfx}").argtypes = [P(OrcGrid), vp, vp, vp, vp]
            getattr(L, f"orc_pressure_fd_{sfx}").argtypes = [P(OrcGrid), vp, vp, vp, vp, OrcSched]
            i64 = ctypes.c_int64
            getattr(L, f"orc_forward_{sfx}").argtypes = [P(OrcGrid), vp, vp, i64, i64, vp, vp, i64, vp, vp]
            getattr(L, f"orc_gradient_{sfx}").argtypes = [P(OrcGrid), vp, vp, ctypes.c_double, i64, i64, vp,
                                                          vp, i64, vp, vp, ctypes.c_int, vp, P(ctypes.c_double)]
            for red in ("sum_abs", "min", "max"):
                f = getattr(L, f"orc_{red}_{sfx}")
                f.argtypes = [P(OrcGrid), vp, P(ctypes.c_int64)]
                f.restype = ctypes.c_double
            for red in ("sum_abs_diff", "max_err", "l1_err"):
                f = getattr(L, f"orc_{red}_{sfx}")
                f.argtypes = [P(OrcGrid), vp, vp, P(ctypes.c_int64)]
                f.restype = ctypes.c_double
        for sfx in ("f32", "f64"):
            f = getattr(L, f"orc_propa_eigen_{sfx}")
            f.argtypes = [P(OrcGrid), P(OrcPropaParams), OrcSched, vp, vp]
            f.restype = ctypes.c_float
        _lib = L
    return _lib


# ---------------------------------------------------------------------------
# Grid layout helpers
# ---------------------------------------------------------------------------
def layout(dim, n, fd_order, storage_fp64=False, nlayer=0, auto_pad=False,
           add_pad=(-1, -1, -1), mul_pad=(-1, -1, -1), offset=(-1, -1, -1)) -> OrcGrid:
    o = OrcLayoutOpts()
    o.dim = dim
    nn = list(n) + [1] * (3 - len(n))
    for a in range(3):
        o.n_inner[a] = nn[a]
        o.add_pad[a] = add_pad[a]
        o.mul_pad[a] = mul_pad[a]
        o.offset[a] = offset[a]
    o.fd_order = fd_order
    o.storage_fp64 = int(storage_fp64)
    o.nlayer = nlayer
    o.auto_pad = int(auto_pad)
    g = OrcGrid()
    if lib().orc_layout(ctypes.byref(o), ctypes.byref(g)) != 0:
        raise ValueError("invalid layout")
    return g


def index_range(g: OrcGrid, ptype: int):
    r = (ctypes.c_int64 * 6)()
    if lib().orc_index_range(ctypes.byref(g), ptype, r) != 0:
        raise ValueError("invalid point type")
    return list(r)


def sched(mode=0, cb=(9999, 4, 16)) -> OrcSched:
    s = OrcSched()
    s.mode, s.cb1, s.cb2, s.cb3 = mode, cb[0], cb[1], cb[2]
    return s


def d2_coefs(fd_order):
    a = (ctypes.c_double * 9)()
    if lib().orc_d2_coefs(fd_order, a) != 0:
        raise ValueError(fd_order)
    return list(a)[: fd_order // 2 + 1]


# ---------------------------------------------------------------------------
# Exact FD weights (pinning src/fdm.h:276-507 from the mathematics)
# ---------------------------------------------------------------------------
def fornberg_d2_weights(R: int):
    """Exact centred second-derivative weights on 2R+1 equispaced nodes.

    Fornberg (1988), 'Generation of finite difference formulas on arbitrarily
    spaced grids', Math. Comp. 51(184): recursion over nodes, exact rationals.
    Returns [A0, A1, ..., AR] (symmetric stencil)."""
    nodes = [0] + [s * j for j in range(1, R + 1) for s in (1, -1)]
    M = 2
    N = len(nodes) - 1
    c = [[[Fraction(0)] * (N + 1) for _ in range(N + 1)] for _ in range(M + 1)]
    c[0][0][0] = Fraction(1)
    c1 = Fraction(1)
    x0 = Fraction(0)
    for n in range(1, N + 1):
        c2 = Fraction(1)
        for nu in range(n):
            c3 = Fraction(nodes[n] - nodes[nu])
            c2 *= c3
            for m in range(min(n, M) + 1):
                prev = c[m - 1][n - 1][nu] if m > 0 else Fraction(0)
                c[m][n][nu] = ((nodes[n] - x0) * c[m][n - 1][nu] - m * prev) / c3
        for m in range(min(n, M) + 1):
            prev = c[m - 1][n - 1][n - 1] if m > 0 else Fraction(0)
            c[m][n][n] = c1 / c2 * (m * prev - (nodes[n - 1] - x0) * c[m][n - 1][n - 1])
        c1 = c2
    w = c[2][N]
    out = [w[0]] + [w[2 * j - 1] for j in range(1, R + 1)]
    for j in range(1, R + 1):
        assert w[2 * j - 1] == w[2 * j]
    return out


# ---------------------------------------------------------------------------
# Propagator scalar arithmetic in float32 (src/propagator_Ac2.cpp:181-226)
# ---------------------------------------------------------------------------
f32 = np.float32


def cfl(fd_order, dim, fp64=False):
    """Propagator_Ac2::getCFL: sqrt(sum|O2| / (dim * sum|Ok|)) stored as Myfloat."""
    a1 = sum_abs_coef(2, fp64)
    a2 = sum_abs_coef(fd_order, fp64) * dim
    v = math.sqrt(a1 / a2)
    return float(v) if fp64 else f32(v)


def sum_abs_coef(fd_order, fp64=False):
    """getSumAbsFD_D2Coef (src/fdm.h:771-787) over Myfloat2 coefficients."""
    a = d2_coefs(fd_order)
    cast = (lambda x: x) if fp64 else (lambda x: float(f32(x)))
    vec = [a[j] for j in range(len(a) - 1, 0, -1)] + [a[0]] + a[1:]
    s = 0.0
    for x in vec:
        s += abs(cast(x))
    return s


def unit_grid(n_inner, inner_start, d_cfg=PI_REF / 30, fp64=False):
    """Grid::defineUnitGrid (src/grid.cpp:1924-1957) for one axis at one rank:
    d = 1/(N-1); origin rescaled then rounded to Myfloat by resetOriginGrid."""
    d = 1.0 / (n_inner - 1)
    orig_old = (0 - inner_start) * d_cfg          # src/grid.cpp:465-468
    orig_new = orig_old * d / d_cfg               # Myfloat Orig1New
    return d, float(orig_new) if fp64 else float(f32(orig_new))


def propa_params(n, fd_order, dim=3, ratio_cfl=1.0, dt_cfg=0.0, snap_dt=0.0, snap_inc=1,
                 tmax_cfg=-1.0, nt_cfg=20, fp64=False):
    """stableDt, dt, snapInc, nt, tmax exactly as Propagator_Ac2::initialize (FP32 Myfloat)."""
    T = (lambda x: float(x)) if fp64 else f32
    n = list(n)
    dmin = T(1.0 / (n[0] - 1))                    # Grid::getMinSpaceSampling returns Myfloat
    for a in range(1, dim):
        dmin = T(min(float(dmin), 1.0 / (n[a] - 1)))
    vmax = T(1.0)
    stable = T(T(cfl(fd_order, dim, fp64) * dmin) / vmax)
    if dt_cfg > 0.0:
        dt = T(dt_cfg)
    else:
        dt = T(float(stable) * ratio_cfl)
    if snap_dt > 0.0:
        inc = int(math.ceil(snap_dt / float(dt)))
        dt = T(snap_dt / inc)
        sinc = int(snap_dt / float(dt))
        if (snap_dt - float(T(T(sinc) * dt))) > float(dt) / 2.0:
            sinc += 1
    else:
        sinc = snap_inc
    if tmax_cfg != -1.0:
        tmax = T(tmax_cfg)
        nt = int(T(T(tmax / dt) + T(1)))
    else:
        nt = nt_cfg
        tmax = T(T(nt - 1) * dt)
    return dict(stable_dt=stable, dt=dt, snap_inc=sinc, nt=nt, tmax=tmax)


def eigen_closed_form(n, fd_order, modes, dim=3, fp64=False, **kw):
    """errTestCase of the eigen-mode Propa test from the discrete recurrence.

    With the free-surface odd extension a sine mode is an eigenvector of the
    FD Laplacian with eigenvalue lam = sum_i S(k_i d_i) inv2_i, where
    S(t) = A0 + 2 sum_j A_j cos(j t) and A_j are the Myfloat2-rounded weights.
    The leapfrog then reduces to A_{n+1} = (2 + c lam) A_n - A_{n-1}, c = v^2 dt^2,
    and the separable spatial factor cancels in errTestCase."""
    p = propa_params(n, fd_order, dim=dim, fp64=fp64, **kw)
    T = (lambda x: float(x)) if fp64 else f32
    dt = p["dt"]
    A = [T(a) for a in d2_coefs(fd_order)]
    lam = 0.0
    for a in range(dim):
        d = 1.0 / (n[a] - 1)
        inv = T(1.0 / d)
        inv2 = float(T(inv * inv))
        th = PI_REF * modes[a] * d
        S = float(A[0]) + 2.0 * sum(float(A[j]) * math.cos(j * th) for j in range(1, len(A)))
        lam += S * inv2
    gamma = math.sqrt(sum(m * m for m in modes[:dim])) if dim > 1 else float(modes[0])
    c = float(T(T(T(1.0) * T(1.0)) * dt) * dt)
    am2 = math.sin(PI_REF * float(T(-2.0 * float(dt))) * gamma)
    am1 = math.sin(PI_REF * float(T(-float(dt))) * gamma)
    s1 = s2 = 0.0
    for it in range(p["nt"]):
        a0 = (2.0 + c * lam) * am1 - am2
        if it % p["snap_inc"] == 0:
            ex = math.sin(PI_REF * float(T(it * dt)) * gamma)
            s1 += abs(a0 - ex)
            s2 += abs(ex)
        am2, am1 = am1, a0
    return s1 / s2, p


# ---------------------------------------------------------------------------
# numpy-facing wrappers over the C kernels
# ---------------------------------------------------------------------------
def _sfx(dtype, compute64=False):
    if dtype == np.float64:
        return "f64"
    return "f32c64" if compute64 else "f32"


def _p(a: np.ndarray):
    assert a.flags.c_contiguous
    return a.ctypes.data_as(ctypes.c_void_p)


def new_grid(g: OrcGrid, dtype=np.float32, fill=0.0):
    return np.full(g.npoint, fill, dtype=dtype)


def fill_sep(g, u, ptype, funcs, params, amp, orig, d, dim):
    r = index_range(g, ptype)
    tabs = []
    for a in range(3):
        lo, hi = r[2 * a], r[2 * a + 1]
        t = np.zeros(hi - lo + 1, dtype=np.float64)
        lib().orc_fill_table(funcs[a], int(a < dim), orig[a], d[a], lo, hi, params[a],
                             t.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        tabs.append(t)
    rr = (ctypes.c_int64 * 6)(*r)
    getattr(lib(), f"orc_fill_sep_{_sfx(u.dtype)}")(ctypes.byref(g), _p(u), rr, _p(tabs[0]),
                                                     _p(tabs[1]), _p(tabs[2]), float(amp))
    return tabs


def fill_const(g, u, ptype, val, compute64=False):
    rr = (ctypes.c_int64 * 6)(*index_range(g, ptype))
    getattr(lib(), f"orc_fill_const_{_sfx(u.dtype, compute64)}")(ctypes.byref(g), _p(u), rr, float(val))


def inv2(d, dtype, compute64=False):
    if dtype == np.float64 or compute64:
        return np.array([lib().orc_inv2_f64(x) if x > 0 else np.inf for x in d], dtype=np.float64)
    return np.array([lib().orc_inv2_f32(x) if x > 0 else np.inf for x in d], dtype=np.float32)


def free_surface(g, u, compute64=False, face_mask=0x3F):
    if face_mask == 0x3F:
        getattr(lib(), f"orc_free_surface_{_sfx(u.dtype, compute64)}")(ctypes.byref(g), _p(u))
    else:
        getattr(lib(), f"orc_free_surface_mask_{_sfx(u.dtype, compute64)}")(ctypes.byref(g), _p(u), face_mask)


def inv1(d, dtype, compute64=False):
    """inv_d = (Myfloat2)(ONE / d) of Grid::FD_D1_N* (src/grid.cpp:3415-3417)."""
    t = np.float64 if (dtype == np.float64 or compute64) else np.float32
    return np.array([t(1.0 / x) if x > 0 else np.inf for x in d], dtype=t)


def d1_coefs(order):
    """staggered D1 weights A1..AR (index 0 unused), as the oracle holds them"""
    a = (ctypes.c_double * 9)()
    if lib().orc_d1_coefs(order, a):
        raise ValueError(order)
    return list(a)


def fd_d1(g, axis, u, w, iv, s=None, compute64=False):
    """Grid::FD_D1_N1/N2/N3: staggered first derivative, w at i - 1/2 on inner points."""
    s = s or sched()
    return getattr(lib(), f"orc_fd_d1_{_sfx(u.dtype, compute64)}")(
        ctypes.byref(g), axis, _p(u), _p(w), _p(iv), s)


def fd_d2(g, axis, u, w, iv, s=None, compute64=False):
    s = s or sched()
    return getattr(lib(), f"orc_fd_d2_{_sfx(u.dtype, compute64)}")(
        ctypes.byref(g), axis, _p(u), _p(w), _p(iv), s)


def laplacian(g, u, w, iv, s=None, compute64=False):
    s = s or sched()
    return getattr(lib(), f"orc_laplacian_{_sfx(u.dtype, compute64)}")(
        ctypes.byref(g), _p(u), _p(w), _p(iv), s)


def update_pressure(g, prn, prc, coef, lap, compute64=False):
    return getattr(lib(), f"orc_update_pressure_{_sfx(prn.dtype, compute64)}")(
        ctypes.byref(g), _p(prn), _p(prc), _p(coef), _p(lap))


def pressure_fd(g, prn, prc, coef, iv, s=None, compute64=False):
    s = s or sched()
    return getattr(lib(), f"orc_pressure_fd_{_sfx(prn.dtype, compute64)}")(
        ctypes.byref(g), _p(prn), _p(prc), _p(coef), _p(iv), s)


def forward(g, coef, iv, nt, src_idx, wavelet, rec_idx, compute64=False):
    """Shot modeling (propagator step + additive point sources), traces[nt, nrec]."""
    src_idx = np.ascontiguousarray(src_idx, np.int64)
    rec_idx = np.ascontiguousarray(rec_idx, np.int64)
    wavelet = np.ascontiguousarray(wavelet, coef.dtype)
    traces = np.zeros((nt, len(rec_idx)), coef.dtype)
    rc = getattr(lib(), f"orc_forward_{_sfx(coef.dtype, compute64)}")(
        ctypes.byref(g), _p(coef), _p(iv), nt, len(src_idx), _p(src_idx), _p(wavelet), len(rec_idx),
        _p(rec_idx), _p(traces))
    if rc:
        raise MemoryError("oracle forward: allocation failed")
    return traces


def gradient(g, coef, iv, dt, nt, src_idx, wavelet, rec_idx, d_obs, compute64=False, residual=False):
    """dJ/dm (double, npoint) and J for J = 1/2 |traces - d_obs|^2, full storage;
    residual=True: d_obs is the residual itself (the adjoint source)."""
    src_idx = np.ascontiguousarray(src_idx, np.int64)
    rec_idx = np.ascontiguousarray(rec_idx, np.int64)
    wavelet = np.ascontiguousarray(wavelet, coef.dtype)
    d_obs = np.ascontiguousarray(d_obs, coef.dtype)
    grad = np.zeros(g.npoint, np.float64)
    mis = ctypes.c_double(0.0)
    rc = getattr(lib(), f"orc_gradient_{_sfx(coef.dtype, compute64)}")(
        ctypes.byref(g), _p(coef), _p(iv), float(dt), nt, len(src_idx), _p(src_idx), _p(wavelet),
        len(rec_idx), _p(rec_idx), _p(d_obs), 1 if residual else 0, _p(grad), ctypes.byref(mis))
    if rc:
        raise MemoryError("oracle gradient: allocation failed")
    return grad, mis.value


def reduce(op, g, u, v=None, ptype=INNER_POINTS):
    rr = (ctypes.c_int64 * 6)(*index_range(g, ptype))
    f = getattr(lib(), f"orc_{op}_{_sfx(u.dtype)}")
    if v is None:
        return f(ctypes.byref(g), _p(u), rr)
    return f(ctypes.byref(g), _p(u), _p(v), rr)


def propa_eigen(n, fd_order, modes, dim=3, fp64=False, split=False, sched_=None, **kw):
    """Run the eigen-mode Propa case on the CPU oracle; returns (errTestCase, prn, grid)."""
    g = layout(dim, n, fd_order, storage_fp64=fp64)
    p = propa_params(n, fd_order, dim=dim, fp64=fp64, **kw)
    P = OrcPropaParams()
    P.dim, P.fd_order, P.nt, P.snap_inc, P.split_comp = dim, fd_order, p["nt"], p["snap_inc"], int(split)
    P.dt = float(p["dt"])
    for a in range(3):
        P.mode[a] = int(modes[a]) if a < len(modes) else 1
        if a < dim:
            dd, oo = unit_grid(n[a], g.ax[a].inner_start, fp64=fp64)
        else:
            dd, oo = 0.0, 0.0
        P.d[a], P.orig[a] = dd, oo
    dtype = np.float64 if fp64 else np.float32
    prn = np.zeros(g.npoint, dtype=dtype)
    prc = np.zeros(g.npoint, dtype=dtype)
    f = getattr(lib(), "orc_propa_eigen_f64" if fp64 else "orc_propa_eigen_f32")
    err = f(ctypes.byref(g), ctypes.byref(P), sched_ or sched(), _p(prn), _p(prc))
    return err, prn, g, p
