"""Acquisition and models for shot modeling: Ricker source, point-source and
receiver-line indices, smooth layered velocity models.

The built-in geometry follows the reference's DataAcquisition
(src/data_Acquisition.cpp:58-305): the source one point below the top n1
boundary (inner i1 = 1) in the middle of n2 x n3, receivers on the plane
i1 = 1; the Ricker wavelet is its appliSourceTerm formula
(src/data_Acquisition.cpp:278-286).  The velocity models are this package's:
a linear increase with depth (n1) times (1 + sigma G), G a smooth seeded
Gaussian random field of unit standard deviation.
"""
from __future__ import annotations

import math

import numpy as np


def ricker(nt, dt, f0, amp=1.0):
    """(1 - 2 a^2) exp(-a^2), a = pi f0 (t - t00), t00 = 1.5 sqrt(6) / (pi f0)"""
    t00 = 1.5 * math.sqrt(6.0) / (math.pi * f0)
    t = np.arange(nt) * dt
    a2 = (math.pi * f0 * (t - t00)) ** 2
    return amp * (1.0 - 2.0 * a2) * np.exp(-a2)


def flat_index(p, i1, i2, i3):
    return int(i1) + int(i2) * p.n1 + int(i3) * p.n1 * p.n2


def source_top_center(p, depth=1):
    """flat index of the point `depth` points below the top inner plane, centre of n2 x n3"""
    r = p.inner()
    return flat_index(p, r[0] + depth, (r[2] + r[3]) // 2, (r[4] + r[5]) // 2)


def receiver_line(p, count=None, depth=1):
    """receivers on the plane `depth` below the top, along n2 at the centre of n3
    (count evenly spaced receivers, default every inner point)"""
    r = p.inner()
    n2 = r[3] - r[2] + 1
    count = n2 if count is None else min(count, n2)
    i2 = np.round(np.linspace(r[2], r[3], count)).astype(np.int64)
    i3 = (r[4] + r[5]) // 2
    return np.array([flat_index(p, r[0] + depth, j, i3) for j in i2], np.int64)


def _linear_resample_matrix(n_in, n_out):
    """(n_out, n_in) weights of linear interpolation with the end points of both
    grids aligned (output point o sits at input coordinate o (n_in-1)/(n_out-1))"""
    x = np.arange(n_out) * ((n_in - 1) / max(1, n_out - 1))
    i0 = np.minimum(np.floor(x).astype(np.int64), n_in - 1)
    i1 = np.minimum(i0 + 1, n_in - 1)
    w = x - i0
    m = np.zeros((n_out, n_in))
    m[np.arange(n_out), i0] += 1.0 - w
    m[np.arange(n_out), i1] += w
    return m


def gaussian_field(shape, seed, corr_points=12.0):
    """smooth Gaussian random field of unit standard deviation: white noise on a
    coarse grid, Gaussian-filtered, trilinearly resampled to `shape` (one axis
    at a time, as three small matrix products)"""
    from scipy import ndimage

    rng = np.random.default_rng(seed)
    step = max(2, int(corr_points // 3))
    coarse = tuple(max(4, s // step + 2) for s in shape)
    f = ndimage.gaussian_filter(rng.standard_normal(coarse), sigma=corr_points / step, mode="wrap")
    m0, m1, m2 = (_linear_resample_matrix(c, s) for c, s in zip(coarse, shape))
    f = np.tensordot(m0, f, axes=(1, 0))                  # (n3, c2, c1)
    f = np.tensordot(m1, f, axes=(1, 1))                  # (n2, n3, c1)
    f = np.ascontiguousarray(f.transpose(1, 0, 2)) @ m2.T  # (n3, n2, n1)
    flat = f.reshape(-1)
    flat -= flat.sum() / flat.size
    sd = math.sqrt(float(np.dot(flat, flat)) / flat.size)
    if sd > 0:
        flat /= sd
    return f


def layered_velocity(p, vmin, vmax, sigma=0.05, seed=0):
    """v over the whole grid ([n3, n2, n1] flattened): vmin at the top inner plane
    to vmax at the bottom along n1, times (1 + sigma G), clipped to [vmin, vmax]"""
    r = p.inner()
    i1 = np.arange(p.n1)
    depth = np.clip((i1 - r[0]) / max(1, r[1] - r[0]), 0.0, 1.0)
    v1 = vmin + (vmax - vmin) * depth
    if sigma <= 0:
        return np.broadcast_to(v1[None, None, :], (p.n3, p.n2, p.n1)).astype(np.float64).ravel()
    v = gaussian_field((p.n3, p.n2, p.n1), seed)    # updated in place: no grid-sized temporaries
    v *= sigma
    v += 1.0
    v *= v1[None, None, :]
    np.clip(v, vmin, vmax, out=v)
    return v.ravel()


def smooth_velocity(p, vmin, vmax, seed=0):
    """smooth seeded velocity in [vmin, vmax] (no layering)"""
    g = gaussian_field((p.n3, p.n2, p.n1), seed, corr_points=max(8.0, min(p.n1, p.n2, p.n3) / 4))
    g = (g - g.min()) / max(g.max() - g.min(), 1e-30)
    return (vmin + (vmax - vmin) * g).ravel()


def stable_dt(p, vmax, cfl_ratio, cfl):
    """dt = cfl_ratio * CFL * min(d) / vmax (Propagator_Ac2::initialize, src/propagator_Ac2.cpp:84-304)"""
    return cfl_ratio * cfl * min(d for d in p.d[:p.dim]) / vmax


def fd_cfl(fd_order, dim=3):
    """sqrt(sum|O2 weights| / (dim sum|weights|)) (src/propagator_Ac2.cpp:410-439), in double"""
    from .plain import FD_D2

    s2 = sum(abs(a) for a in FD_D2[2][:1]) + 2 * sum(abs(a) for a in FD_D2[2][1:])
    sk = abs(FD_D2[fd_order][0]) + 2 * sum(abs(a) for a in FD_D2[fd_order][1:])
    return math.sqrt(s2 / (dim * sk))
