"""Grid index layout of the hpcscan Grid class, for Python callers of the C ABI.

Restates Grid::initializeGrid for one rank (src/grid.cpp:252-468), with the
padding and offset rules of src/grid.cpp:839-1009:

    Offset | Halo1 | Layer1 | Inner | Layer2 | Halo2 | Pad     (per axis)

and the unit-grid sampling of Grid::defineUnitGrid (src/grid.cpp:1924-1957).
The C++ host harness has its own copy of this logic (hpcscan_amd/host/grid.cpp);
tests check both against the oracle.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import List, Sequence

import numpy as np

UNSPEC = -1


@dataclass
class Axis:
    offset_start: int = 0
    offset_end: int = 0
    halo1_start: int = 0
    halo1_end: int = 0
    layer1_start: int = 0
    layer1_end: int = 0
    inner_start: int = 0
    inner_end: int = 0
    layer2_start: int = 0
    layer2_end: int = 0
    halo2_start: int = 0
    halo2_end: int = 0
    pad_start: int = 0
    pad_end: int = 0
    n: int = 1


@dataclass
class GridLayout:
    dim: int
    fd_order: int
    fp64: bool
    axes: List[Axis] = field(default_factory=list)
    d: List[float] = field(default_factory=lambda: [0.0, 0.0, 0.0])

    @property
    def halo(self) -> int:
        return self.fd_order // 2

    @property
    def n1(self) -> int:
        return self.axes[0].n

    @property
    def n2(self) -> int:
        return self.axes[1].n

    @property
    def n3(self) -> int:
        return self.axes[2].n

    @property
    def npoint(self) -> int:
        return self.n1 * self.n2 * self.n3

    @property
    def dtype(self):
        return np.float64 if self.fp64 else np.float32

    def inner(self) -> List[int]:
        return [v for a in self.axes for v in (a.inner_start, a.inner_end)]

    def all_points(self) -> List[int]:
        return [v for a in self.axes for v in (a.halo1_start, a.halo2_end)]

    def n_inner_points(self) -> int:
        r = self.inner()
        return (r[1] - r[0] + 1) * (r[3] - r[2] + 1) * (r[5] - r[4] + 1)

    def shape3(self):
        """numpy view shape (n3, n2, n1) of the flat array."""
        return (self.n3, self.n2, self.n1)


def _axis(a: int, n_inner: int, h: int, fp64: bool, nlayer: int, auto_pad: bool,
          add_pad: int, mul_pad: int, offset: int) -> Axis:
    x = Axis()
    off = offset
    if off == UNSPEC:
        off = 1 if (a == 0 and h % 2 == 1 and not fp64) else 0
    x.offset_start = 0
    x.offset_end = off - 1
    x.halo1_start = x.offset_end + 1
    x.halo1_end = x.halo1_start + h - 1
    x.layer1_start = x.halo1_end + 1
    x.layer1_end = x.layer1_start if nlayer == 0 else x.layer1_start + nlayer - 1
    x.inner_start = x.layer1_start if nlayer == 0 else x.layer1_end + 1
    x.inner_end = x.inner_start + n_inner - 1
    x.layer2_start = x.inner_end if nlayer == 0 else x.inner_end + 1
    x.layer2_end = x.inner_end if nlayer == 0 else x.layer2_start + nlayer - 1
    x.halo2_start = x.layer2_end + 1
    x.halo2_end = x.halo2_start + h - 1
    if auto_pad:
        x.pad_start = x.pad_end = x.halo2_end
        if a == 0 and (x.halo2_end + 1) % 2 and not fp64:
            x.pad_start = x.pad_end = x.halo2_end + 1
    elif add_pad != UNSPEC:
        x.pad_start = x.halo2_end + 1
        x.pad_end = x.pad_start + add_pad - 1
    elif mul_pad != UNSPEC:
        x.pad_start = x.halo2_end + 1
        ntmp = (x.pad_start + 1) // mul_pad
        ntot = (ntmp + 1) * mul_pad if mul_pad * ntmp < x.pad_start else ntmp * mul_pad
        x.pad_end = ntot - 1
    else:
        x.pad_start = x.pad_end = x.halo2_end
    x.n = x.pad_end + 1
    return x


def make_layout(dim: int, n: Sequence[int], fd_order: int, fp64: bool = False, nlayer: int = 0,
                auto_pad: bool = False, add_pad=(UNSPEC,) * 3, mul_pad=(UNSPEC,) * 3,
                offset=(UNSPEC,) * 3, unit_grid: bool = True) -> GridLayout:
    if dim not in (1, 2, 3):
        raise ValueError(f"dim {dim}")
    if fd_order not in (2, 4, 6, 8, 10, 12, 14, 16):
        raise ValueError(f"fd_order {fd_order}")
    n = list(n) + [1] * (3 - len(n))
    h = fd_order // 2
    g = GridLayout(dim=dim, fd_order=fd_order, fp64=fp64)
    for a in range(3):
        if a < dim:
            g.axes.append(_axis(a, n[a], h, fp64, nlayer, auto_pad, add_pad[a], mul_pad[a], offset[a]))
        else:
            g.axes.append(Axis())
    if unit_grid:
        g.d = [1.0 / (n[a] - 1) if a < dim else 0.0 for a in range(3)]
    return g


def unit_origin(g: GridLayout, a: int, d_cfg: float = 3.1415926535897 / 30) -> float:
    """Origin of axis a after defineUnitGrid: rescaled, then rounded to Myfloat
    (src/grid.cpp:1944-1951, resetOriginGrid takes Myfloat)."""
    if a >= g.dim:
        return 0.0
    orig_old = (0 - g.axes[a].inner_start) * d_cfg
    v = orig_old * g.d[a] / d_cfg
    return float(v) if g.fp64 else float(np.float32(v))
