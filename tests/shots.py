"""Synthetic shots for the modeling / adjoint tests: smooth seeded velocity on
the unit cube, a Ricker point source one point below the top (n1) boundary
plane, a line of receivers on the plane below the boundary, observed data from
a smoothly perturbed model.  Mirrors the reference's built-in acquisition
(src/data_Acquisition.cpp:58-305: source at i1 = 1 in the middle of n2 x n3,
receivers at i1 = 1) with an additive source."""
import math

import numpy as np

from hpcscan_amd.layout import make_layout


def ricker(nt, dt, f0, amp=1.0):
    """Ricker wavelet of peak frequency f0 delayed by 1.5 sqrt(6) / (pi f0), as
    DataAcquisition::appliSourceTerm (src/data_Acquisition.cpp:278-286)."""
    da = math.pi * f0
    t00 = 1.5 * math.sqrt(6.0) / da
    t = np.arange(nt) * dt
    a2 = (math.pi * f0 * (t - t00)) ** 2
    return amp * (1.0 - 2.0 * a2) * np.exp(-a2)


def smooth_field(p, seed, nmodes=4):
    """sum of a few seeded low-wavenumber sines over the whole grid, in [-1, 1]"""
    rng = np.random.default_rng(seed)
    n1, n2, n3 = p.n1, p.n2, p.n3
    z, y, x = np.meshgrid(np.linspace(0, 1, n3), np.linspace(0, 1, n2), np.linspace(0, 1, n1), indexing="ij")
    f = np.zeros((n3, n2, n1))
    for _ in range(nmodes):
        k = rng.uniform(0.5, 2.0, 3)
        ph = rng.uniform(0, 2 * math.pi, 3)
        f += np.sin(k[0] * math.pi * x + ph[0]) * np.sin(k[1] * math.pi * y + ph[1]) * np.sin(
            k[2] * math.pi * z + ph[2])
    return (f / nmodes).ravel()


def masks(p):
    inner = np.zeros((p.n3, p.n2, p.n1), bool)
    r = p.inner()
    inner[r[4]:r[5] + 1, r[2]:r[3] + 1, r[0]:r[1] + 1] = True
    interior = np.zeros_like(inner)
    interior[r[4] + 1:r[5], r[2] + 1:r[3], r[0] + 1:r[1]] = True
    return inner.ravel(), interior.ravel()


def make_shot(oracle, n_inner, order, nt, fp64=False, seed=0, vmin=1.0, vmax=2.0, ppw=8.0, cfl_ratio=0.6,
              perturb=0.03, with_obs=True):
    dtype = np.float64 if fp64 else np.float32
    p = make_layout(3, n_inner, order, fp64=fp64)
    o = oracle.layout(3, list(n_inner), order, storage_fp64=fp64)
    assert o.npoint == p.npoint
    h = min(p.d)
    v = 0.5 * (vmin + vmax) + 0.5 * (vmax - vmin) * smooth_field(p, seed)
    m = 1.0 / (v * v)
    cfl = oracle.cfl(order, 3)
    dt = float(dtype(cfl_ratio * cfl * h / vmax))
    inner, interior = masks(p)
    coef = np.where(inner, dt * dt / m, 0.0).astype(dtype)
    r = p.inner()
    mid2, mid3 = (r[2] + r[3]) // 2, (r[4] + r[5]) // 2
    src = np.array([(r[0] + 2) + mid2 * p.n1 + mid3 * p.n1 * p.n2], np.int64)
    rec = np.array([(r[0] + 1) + i2 * p.n1 + mid3 * p.n1 * p.n2 for i2 in range(r[2] + 1, r[3])], np.int64)
    f0 = vmin / (2.5 * ppw * h)
    wavelet = (ricker(nt, dt, f0) * (dt * dt) * 1e3).astype(dtype).reshape(nt, 1)
    iv = oracle.inv2(p.d, dtype)
    shot = dict(p=p, o=o, m=m, coef=coef, dt=dt, nt=nt, src=src, rec=rec, wavelet=wavelet, iv=iv,
                inner_mask=inner, grad_mask=interior, dtype=dtype)
    if with_obs:
        # observed data: the same shot in the perturbed model
        m_true = m * (1.0 + perturb * smooth_field(p, seed + 100, nmodes=3))
        coef_true = np.where(inner, dt * dt / m_true, 0.0).astype(dtype)
        shot["d_obs"] = oracle.forward(o, coef_true, iv, nt, src, wavelet, rec)
        shot["coef_true"] = coef_true
    return shot
