"""torch.autograd wrapper of shot modeling, for tests: the forward is
hsx_forward, the vector-Jacobian product is hsx_gradient (the hand-written
adjoint) with the incoming trace gradient passed as the residual itself
(HSX_ADJ_RESIDUAL), so that a loss of any scale or weighting reaches it
unrounded.

Device arrays are torch tensors on the GPU; the library is called through the
C ABI (hpcscan_amd.capi), which fails loudly when it is not built.
"""
from __future__ import annotations

import torch

from . import capi


class Ac2Shot:
    """Geometry of one shot on one grid: source / receiver indices and wavelet
    as device tensors, the descriptor, the mask of the inner points."""

    def __init__(self, layout, dt, src_idx, wavelet, rec_idx, device="cuda"):
        self.layout = layout
        self.g = capi.hsx_grid(layout)
        self.dtype = torch.float64 if layout.fp64 else torch.float32
        self.dt = float(dt)
        self.src = torch.as_tensor(src_idx, dtype=torch.int64, device=device).contiguous()
        self.rec = torch.as_tensor(rec_idx, dtype=torch.int64, device=device).contiguous()
        self.wavelet = torch.as_tensor(wavelet, dtype=self.dtype, device=device).contiguous()
        self.nt = int(self.wavelet.shape[0])
        self.nrec = int(self.rec.numel())
        self.desc = capi.make_shot(self.nt, self.dt, (self.src, int(self.src.numel())), self.wavelet,
                                   (self.rec, self.nrec), npoint=layout.npoint, host_src=src_idx, host_rec=rec_idx)
        r = layout.inner()
        mask = torch.zeros((layout.n3, layout.n2, layout.n1), dtype=torch.bool, device=device)
        mask[r[4]:r[5] + 1, r[2]:r[3] + 1, r[0]:r[1] + 1] = True
        self.inner = mask.reshape(-1)

    def coef(self, m: torch.Tensor) -> torch.Tensor:
        """coef = dt^2 / m on the inner points, 0 elsewhere (storage type)"""
        c = torch.where(self.inner, (self.dt * self.dt) / m, torch.zeros_like(m))
        return c.to(self.dtype).contiguous()


class Ac2Modeling(torch.autograd.Function):
    """traces = F(m) for squared slowness m (npoint values); backward = adjoint."""

    @staticmethod
    def forward(ctx, m, shot: Ac2Shot, ckpt_every: int = 0, ckpt_format: int = capi.HSX_CKPT_FP32):
        coef = shot.coef(m.detach())
        traces = torch.empty((shot.nt, shot.nrec), dtype=shot.dtype, device=m.device)
        torch.cuda.synchronize(m.device)          # inputs written on torch's stream
        capi.forward(shot.g, coef, shot.desc, traces)
        ctx.save_for_backward(coef, traces)
        ctx.shot, ctx.every, ctx.fmt = shot, ckpt_every, ckpt_format
        ctx.m_dtype = m.dtype
        return traces

    @staticmethod
    def backward(ctx, grad_traces):
        coef, traces = ctx.saved_tensors
        shot = ctx.shot
        # the incoming trace gradient is the adjoint source itself (HSX_ADJ_RESIDUAL)
        res = grad_traces.to(shot.dtype).contiguous()
        grad = torch.zeros(coef.numel(), dtype=torch.float64, device=coef.device)
        torch.cuda.synchronize(coef.device)
        capi.gradient(shot.g, coef, shot.desc, res, grad, ctx.every, ctx.fmt, residual=True)
        return grad.to(ctx.m_dtype), None, None, None


def modeling(m: torch.Tensor, shot: Ac2Shot, ckpt_every: int = 0, ckpt_format: int = capi.HSX_CKPT_FP32):
    return Ac2Modeling.apply(m, shot, ckpt_every, ckpt_format)
