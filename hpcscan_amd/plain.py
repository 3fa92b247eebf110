"""Plain PyTorch version of the shot modeling and its gradient, for comparison
only: the same discrete scheme written with tensor operations, FP32 (or FP64),
every time step kept on the autograd tape (full storage), the gradient from
torch.autograd.  It calls no kernel of this library; bench.py times it beside
the hand-written adjoint, and a GPU test checks that both give the same
gradient.

Scheme (as hsx_forward / hsx_gradient): q^{n+1} = B(2 q^n - q^{n-1} + c Lap(q^n)
+ f^n) with B the free surface (src/grid.cpp:3065-3230), Lap the FD Laplacian
of the reference's coefficient table (src/fdm.h:276-507), c = dt^2 / m.
"""
from __future__ import annotations

import torch

FD_D2 = {
    2: [-2.0, 1.0],
    4: [-5 / 2, 4 / 3, -1 / 12],
    6: [-49 / 18, 3 / 2, -3 / 20, 1 / 90],
    8: [-205 / 72, 8 / 5, -1 / 5, 8 / 315, -1 / 560],
    10: [-5269 / 1800, 5 / 3, -5 / 21, 5 / 126, -5 / 1008, 1 / 3150],
    12: [-5369 / 1800, 12 / 7, -15 / 56, 10 / 189, -1 / 112, 2 / 1925, -1 / 16632],
    14: [-266681 / 88200, 7 / 4, -7 / 24, 7 / 108, -7 / 528, 7 / 3300, -7 / 30888, 1 / 84084],
    16: [-1077749 / 352800, 16 / 9, -14 / 45, 112 / 1485, -7 / 396, 112 / 32175, -2 / 3861, 16 / 315315,
         -1 / 411840],
}


class PlainModel:
    """Layout-aware torch operators on [n3, n2, n1] tensors of one grid."""

    def __init__(self, layout, dt):
        self.p = layout
        self.R = layout.fd_order // 2
        r = layout.inner()
        self.s = (r[4], r[2], r[0])          # inner start per tensor dim (i3, i2, i1)
        self.e = (r[5], r[3], r[1])
        self.A = FD_D2[layout.fd_order]
        self.inv2 = [1.0 / (d * d) for d in (layout.d[2], layout.d[1], layout.d[0])]
        self.dt = dt

    def _sl(self, dim, idx):
        """index tuple: idx on dim, inner range on the two other dims"""
        out = []
        for k in range(3):
            out.append(idx if k == dim else slice(self.s[k], self.e[k] + 1))
        return tuple(out)

    def free_surface(self, u):
        """the six phases in the reference order: axis n1, n2, n3 (tensor dims 2, 1, 0)"""
        u = u.clone()
        R = self.R
        for dim in (2, 1, 0):
            b0, b1 = self.s[dim], self.e[dim]
            u[self._sl(dim, b0)] = 0
            u[self._sl(dim, slice(b0 - R, b0))] = -torch.flip(u[self._sl(dim, slice(b0 + 1, b0 + R + 1))], [dim])
            u[self._sl(dim, b1)] = 0
            u[self._sl(dim, slice(b1 + 1, b1 + R + 1))] = -torch.flip(u[self._sl(dim, slice(b1 - R, b1))], [dim])
        return u

    def laplacian(self, u):
        """on the inner box, from a field with halos"""
        R, s, e = self.R, self.s, self.e
        c = u[s[0]:e[0] + 1, s[1]:e[1] + 1, s[2]:e[2] + 1]
        lap = torch.zeros_like(c)
        for dim in range(3):
            acc = self.A[0] * c
            for j in range(1, R + 1):
                sl_p = [slice(s[k], e[k] + 1) for k in range(3)]
                sl_m = list(sl_p)
                sl_p[dim] = slice(s[dim] + j, e[dim] + j + 1)
                sl_m[dim] = slice(s[dim] - j, e[dim] - j + 1)
                acc = acc + self.A[j] * (u[tuple(sl_p)] + u[tuple(sl_m)])
            lap = lap + acc * self.inv2[dim]
        return lap

    def step(self, prev, cur, coef_inner):
        """next wavefield (new tensor) from B(cur) and prev"""
        q = self.free_surface(cur)
        s, e = self.s, self.e
        box = (slice(s[0], e[0] + 1), slice(s[1], e[1] + 1), slice(s[2], e[2] + 1))
        nxt = torch.zeros_like(q)
        nxt[box] = 2.0 * q[box] - prev[box] + coef_inner * self.laplacian(q)
        return q, nxt


def plain_forward(layout, m, dt, src_idx, wavelet, rec_idx):
    """traces [nt, nrec] of the shot for squared slowness m (flat tensor with
    autograd), storage dtype of m"""
    pm = PlainModel(layout, dt)
    shape = (layout.n3, layout.n2, layout.n1)
    s, e = pm.s, pm.e
    box = (slice(s[0], e[0] + 1), slice(s[1], e[1] + 1), slice(s[2], e[2] + 1))
    coef_inner = (dt * dt) / m.reshape(shape)[box]
    prev = torch.zeros(shape, dtype=m.dtype, device=m.device)
    cur = torch.zeros(shape, dtype=m.dtype, device=m.device)
    src_idx = torch.as_tensor(src_idx, device=m.device)
    rec_idx = torch.as_tensor(rec_idx, device=m.device)
    wavelet = torch.as_tensor(wavelet, dtype=m.dtype, device=m.device).reshape(-1, src_idx.numel())
    traces = []
    for n in range(wavelet.shape[0]):
        q, nxt = pm.step(prev, cur, coef_inner)
        flat = nxt.reshape(-1)
        flat = flat.index_add(0, src_idx, wavelet[n])
        traces.append(flat[rec_idx])
        prev, cur = q, flat.reshape(shape)       # the previous field keeps its free surface
    return torch.stack(traces)
