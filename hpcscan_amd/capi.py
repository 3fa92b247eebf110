"""ctypes binding of libhpcscan_mi355x.so (the C ABI of include/hpcscan_mi355x.h).

Used by the tests, bench.py and the autograd harness.  The library is built
in-tree (hpcscan_amd/lib) by ``__graft_entry__.build()``; loading fails loudly
when it is missing -- there is no host fallback for any entry point.

PyTorch, when importable, is imported before the library is loaded so that
both resolve the same HIP runtime (torch ships its own libamdhip64.so.7).
"""
from __future__ import annotations

import ctypes
import os
from typing import Optional, Sequence

import numpy as np

from .layout import GridLayout

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libhpcscan_mi355x.so")

HSX_OK = 0
HSX_ERR_INVALID, HSX_ERR_HIP, HSX_ERR_NOMEM, HSX_ERR_UNSUPPORTED = -1, -2, -3, -4
HSX_FP32, HSX_FP64 = 0, 1
HSX_KERNEL_AUTO, HSX_KERNEL_REF, HSX_KERNEL_TILED = 0, 1, 2
HSX_TILE_CACHED, HSX_TILE_NT, HSX_TILE_NO_RING = 1, 2, 4
HSX_ARRAY_FILL, HSX_ARRAY_COPY, HSX_ARRAY_ADD, HSX_ARRAY_MULTIPLY, HSX_ARRAY_ADDTO = range(5)
ALL_FACES = 0x3F

try:  # one HIP runtime per process: let torch load it first
    import torch  # noqa: F401
except Exception:  # pragma: no cover - torch is optional for the C ABI
    torch = None


class HsxError(RuntimeError):
    def __init__(self, rc: int, msg: str):
        super().__init__(f"hsx error {rc}: {msg}")
        self.rc = rc


class HsxGrid(ctypes.Structure):
    _fields_ = [("dim", ctypes.c_int32), ("fd_order", ctypes.c_int32), ("storage", ctypes.c_int32),
                ("compute", ctypes.c_int32), ("n1", ctypes.c_int64), ("n2", ctypes.c_int64),
                ("n3", ctypes.c_int64), ("inner", ctypes.c_int64 * 6), ("halo1", ctypes.c_int64 * 6),
                ("halo2", ctypes.c_int64 * 6), ("d", ctypes.c_double * 3)]


class HsxDeviceProps(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char * 128), ("arch", ctypes.c_char * 64),
                ("compute_units", ctypes.c_int32), ("clock_mhz", ctypes.c_int32),
                ("mem_clock_mhz", ctypes.c_int32), ("bus_width_bits", ctypes.c_int32),
                ("hbm_bytes", ctypes.c_int64), ("lds_bytes_per_block", ctypes.c_int32),
                ("l2_bytes", ctypes.c_int32)]


# every entry point declared in include/hpcscan_mi355x.h: name -> (restype, argtypes)
_vp, _i, _i64, _d, _sz = ctypes.c_void_p, ctypes.c_int, ctypes.c_int64, ctypes.c_double, ctypes.c_size_t
_P = ctypes.POINTER
_G = _P(HsxGrid)
_R = _P(ctypes.c_int64)
SIGNATURES = {
    "hsx_abi_version": (_i, []),
    "hsx_last_error": (ctypes.c_char_p, []),
    "hsx_init": (_i, [_i]),
    "hsx_device_count": (_i, [_P(_i)]),
    "hsx_device_info": (_i, [_i, _P(HsxDeviceProps)]),
    "hsx_malloc": (_i, [_P(_vp), _sz]),
    "hsx_free": (_i, [_vp]),
    "hsx_memcpy_h2d": (_i, [_vp, _vp, _sz, _vp]),
    "hsx_memcpy_d2h": (_i, [_vp, _vp, _sz, _vp]),
    "hsx_memcpy_d2d": (_i, [_vp, _vp, _sz, _vp]),
    "hsx_stream_create": (_i, [_P(_vp)]),
    "hsx_stream_destroy": (_i, [_vp]),
    "hsx_stream_sync": (_i, [_vp]),
    "hsx_event_create": (_i, [_P(_vp)]),
    "hsx_event_destroy": (_i, [_vp]),
    "hsx_event_record": (_i, [_vp, _vp]),
    "hsx_event_elapsed_ms": (_i, [_vp, _vp, _P(ctypes.c_float)]),
    "hsx_set_kernel": (_i, [_i]),
    "hsx_get_kernel": (_i, []),
    "hsx_set_tile": (_i, [_i, _i, _i, _i]),
    "hsx_set_contract": (_i, [_i]),
    "hsx_get_contract": (_i, []),
    "hsx_fill_const": (_i, [_G, _vp, _R, _d, _vp]),
    "hsx_fill_separable": (_i, [_G, _vp, _R, _P(_d), _P(_d), _P(_d), _d, _vp]),
    "hsx_array_op": (_i, [_G, _i, _vp, _vp, _vp, _d, _vp]),
    "hsx_free_surface": (_i, [_G, _vp, _i, _vp]),
    "hsx_fd_d2": (_i, [_G, _i, _vp, _vp, _vp]),
    "hsx_laplacian": (_i, [_G, _vp, _vp, _vp]),
    "hsx_update_pressure": (_i, [_G, _vp, _vp, _vp, _vp, _vp]),
    "hsx_ac2_step": (_i, [_G, _vp, _vp, _vp, _vp]),
    "hsx_sum_abs": (_i, [_G, _vp, _R, _P(_d), _vp]),
    "hsx_sum_abs_diff": (_i, [_G, _vp, _vp, _R, _P(_d), _vp]),
    "hsx_l1_sums": (_i, [_G, _vp, _vp, _R, _P(_d), _vp]),
    "hsx_l1_sums_noalloc": (_i, [_G, _vp, _vp, _R, _vp, _vp]),
    "hsx_max_err": (_i, [_G, _vp, _vp, _R, _d, _P(_d), _vp]),
    "hsx_min_max": (_i, [_G, _vp, _R, _P(_d), _vp]),
    "hsx_ac2_step_to": (_i, [_G, _vp, _vp, _vp, _vp, _vp]),
    "hsx_ac2_step_fs": (_i, [_G, _vp, _vp, _vp, _i, _i, _vp]),
    "hsx_fd_d1": (_i, [_G, _i, _vp, _vp, _vp]),
    "hsx_inject": (_i, [_G, _vp, _i64, _vp, _vp, _vp, _d, _i, _vp]),
    "hsx_gather": (_i, [_G, _vp, _i64, _vp, _vp, _vp]),
    "hsx_imaging": (_i, [_G, _vp, _vp, _vp, _vp, _vp, _R, _d, _i64, _d, _vp]),
    "hsx_ckpt_bytes": (_sz, [_i64]),
    "hsx_ckpt_encode": (_i, [_vp, _vp, _i64, _vp, _i, _vp]),
    "hsx_ckpt_decode": (_i, [_vp, _i64, _i, _vp, _vp, _vp]),
    "hsx_forward": (_i, [_G, _vp, _vp, _vp, _vp]),
    "hsx_gradient": (_i, [_G, _vp, _vp, _vp, _vp, _vp, _P(_d), _vp]),
    "hsx_adjoint_plan_create": (_i, [_G, _i64, _i64, _vp, _P(_vp)]),
    "hsx_gradient_plan": (_i, [_vp, _G, _vp, _vp, _vp, _vp, _P(_d), _vp]),
    "hsx_adjoint_plan_destroy": (_i, [_vp]),
    "hsx_autotune": (_i, [_G, _i, _P(_i), _P(ctypes.c_float)]),
    "hsx_autotune_clear": (_i, []),
    "hsx_batch_create": (_i, [_i, _G, _P(_vp), _vp, _P(_vp), _P(_vp)]),
    "hsx_batch_run": (_i, [_vp, _i, _vp]),
    "hsx_batch_autotune": (_i, [_vp, _i, _P(_i), _P(ctypes.c_float)]),
    "hsx_batch_set_shape": (_i, [_vp, _i, _i, _i, _i]),
    "hsx_batch_destroy": (_i, [_vp]),
}


class HsxShot(ctypes.Structure):
    _fields_ = [("nt", ctypes.c_int64), ("dt", ctypes.c_double), ("nsrc", ctypes.c_int64),
                ("src_idx", ctypes.c_void_p), ("wavelet", ctypes.c_void_p), ("nrec", ctypes.c_int64),
                ("rec_idx", ctypes.c_void_p)]


class HsxAdjointOpts(ctypes.Structure):
    _fields_ = [("ckpt_every", ctypes.c_int), ("ckpt_format", ctypes.c_int), ("flags", ctypes.c_int)]


HSX_ADJ_RESIDUAL, HSX_ADJ_AUTOTUNE = 1, 2


HSX_INJECT_ADD, HSX_INJECT_SET = 0, 1
HSX_FOLD_PRC, HSX_FOLD_PRN, HSX_FOLD_HALO = 1, 2, 4
HSX_CKPT_FP32, HSX_CKPT_FP16, HSX_CKPT_BF16 = 0, 1, 2

_lib = None


def lib():
    """Load the library once; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with __graft_entry__.build()")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def last_error() -> str:
    return (lib().hsx_last_error() or b"").decode(errors="replace")


def check(rc: int):
    if rc != HSX_OK:
        raise HsxError(rc, last_error())


def init(device: int = 0):
    check(lib().hsx_init(device))


def device_count() -> int:
    n = ctypes.c_int(0)
    check(lib().hsx_device_count(ctypes.byref(n)))
    return n.value


def device_info(device: int = 0) -> dict:
    p = HsxDeviceProps()
    check(lib().hsx_device_info(device, ctypes.byref(p)))
    return {"name": p.name.decode(), "arch": p.arch.decode(), "compute_units": p.compute_units,
            "clock_mhz": p.clock_mhz, "hbm_bytes": p.hbm_bytes, "l2_bytes": p.l2_bytes,
            "lds_bytes_per_block": p.lds_bytes_per_block}


def set_kernel(family: int):
    check(lib().hsx_set_kernel(family))


def set_tile(lanes_n1: int = 0, chunk_planes: int = 0, rows_n2: int = 0, flags: int = 0):
    """Tile shape of the tiled kernels; 0 = the library's choice for each.
    flags: HSX_TILE_CACHED or HSX_TILE_NT (streams), | HSX_TILE_NO_RING."""
    check(lib().hsx_set_tile(lanes_n1, rows_n2, chunk_planes, flags))


def set_contract(on: bool):
    """hsx_set_contract: contracted (FMA) tiled fused step and Laplacian, opt-in."""
    check(lib().hsx_set_contract(1 if on else 0))


def autotune(g: "HsxGrid", reps: int = 3):
    """Pick the fastest tile shape for g's layout; returns ((lanes, rows, chunk, flags), ms)."""
    shape = (ctypes.c_int * 4)()
    ms = ctypes.c_float(0.0)
    check(lib().hsx_autotune(ctypes.byref(g), int(reps), shape, ctypes.byref(ms)))
    return tuple(shape), ms.value


def autotune_clear():
    check(lib().hsx_autotune_clear())


def hsx_grid(g: GridLayout, compute_fp64: Optional[bool] = None) -> HsxGrid:
    """C descriptor of a layout (storage from g.fp64, compute defaults to storage)."""
    c = HsxGrid()
    c.dim, c.fd_order = g.dim, g.fd_order
    c.storage = HSX_FP64 if g.fp64 else HSX_FP32
    cf = g.fp64 if compute_fp64 is None else compute_fp64
    c.compute = HSX_FP64 if cf else HSX_FP32
    c.n1, c.n2, c.n3 = g.n1, g.n2, g.n3
    for a, x in enumerate(g.axes):
        c.inner[2 * a], c.inner[2 * a + 1] = x.inner_start, x.inner_end
        c.halo1[2 * a], c.halo1[2 * a + 1] = x.halo1_start, x.halo1_end
        c.halo2[2 * a], c.halo2[2 * a + 1] = x.halo2_start, x.halo2_end
        c.d[a] = g.d[a]
    return c


def rng(r: Sequence[int]):
    return (ctypes.c_int64 * 6)(*[int(v) for v in r])


class DeviceArray:
    """A device buffer owned by the library allocator."""

    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        p = ctypes.c_void_p()
        check(lib().hsx_malloc(ctypes.byref(p), self.nbytes))
        self.ptr = p

    @classmethod
    def from_numpy(cls, a: np.ndarray) -> "DeviceArray":
        a = np.ascontiguousarray(a)
        d = cls(a.nbytes)
        d.upload(a)
        return d

    def upload(self, a: np.ndarray, stream=None):
        a = np.ascontiguousarray(a)
        assert a.nbytes == self.nbytes
        check(lib().hsx_memcpy_h2d(self.ptr, a.ctypes.data_as(ctypes.c_void_p), a.nbytes, stream))

    def download(self, dtype, count: Optional[int] = None, stream=None) -> np.ndarray:
        count = self.nbytes // np.dtype(dtype).itemsize if count is None else count
        out = np.empty(count, dtype=dtype)
        check(lib().hsx_memcpy_d2h(out.ctypes.data_as(ctypes.c_void_p), self.ptr, out.nbytes, stream))
        return out

    def free(self):
        if self.ptr is not None and self.ptr.value:
            lib().hsx_free(self.ptr)
        self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def dptr(x) -> ctypes.c_void_p:
    if isinstance(x, DeviceArray):
        return x.ptr
    if torch is not None and isinstance(x, torch.Tensor):
        return ctypes.c_void_p(x.data_ptr())
    if x is None:
        return None
    return ctypes.c_void_p(int(x))


# ---- thin wrappers --------------------------------------------------------------
def fill_const(g: HsxGrid, u, r, val, stream=None):
    check(lib().hsx_fill_const(ctypes.byref(g), dptr(u), rng(r), float(val), stream))


def fill_separable(g: HsxGrid, u, r, t1, t2, t3, amp, stream=None):
    t = [np.ascontiguousarray(x, dtype=np.float64) for x in (t1, t2, t3)]
    ptrs = [x.ctypes.data_as(ctypes.POINTER(ctypes.c_double)) for x in t]
    check(lib().hsx_fill_separable(ctypes.byref(g), dptr(u), rng(r), *ptrs, float(amp), stream))


def array_op(g: HsxGrid, op, w, u=None, v=None, val=0.0, stream=None):
    check(lib().hsx_array_op(ctypes.byref(g), op, dptr(w), dptr(u), dptr(v), float(val), stream))


def free_surface(g: HsxGrid, u, face_mask=ALL_FACES, stream=None):
    check(lib().hsx_free_surface(ctypes.byref(g), dptr(u), face_mask, stream))


def fd_d2(g: HsxGrid, axis, u, w, stream=None):
    check(lib().hsx_fd_d2(ctypes.byref(g), axis, dptr(u), dptr(w), stream))


def fd_d1(g: HsxGrid, axis, u, w, stream=None):
    check(lib().hsx_fd_d1(ctypes.byref(g), axis, dptr(u), dptr(w), stream))


def laplacian(g: HsxGrid, u, w, stream=None):
    check(lib().hsx_laplacian(ctypes.byref(g), dptr(u), dptr(w), stream))


def update_pressure(g: HsxGrid, prn, prc, coef, lapla, stream=None):
    check(lib().hsx_update_pressure(ctypes.byref(g), dptr(prn), dptr(prc), dptr(coef), dptr(lapla), stream))


def ac2_step(g: HsxGrid, prn, prc, coef, stream=None):
    check(lib().hsx_ac2_step(ctypes.byref(g), dptr(prn), dptr(prc), dptr(coef), stream))


def sum_abs(g: HsxGrid, u, r, stream=None) -> float:
    out = ctypes.c_double()
    check(lib().hsx_sum_abs(ctypes.byref(g), dptr(u), rng(r), ctypes.byref(out), stream))
    return out.value


def sum_abs_diff(g: HsxGrid, u, v, r, stream=None) -> float:
    out = ctypes.c_double()
    check(lib().hsx_sum_abs_diff(ctypes.byref(g), dptr(u), dptr(v), rng(r), ctypes.byref(out), stream))
    return out.value


def l1_sums(g: HsxGrid, u, v, r, stream=None):
    out = (ctypes.c_double * 2)()
    check(lib().hsx_l1_sums(ctypes.byref(g), dptr(u), dptr(v), rng(r), out, stream))
    return out[0], out[1]


def max_err(g: HsxGrid, u, v, r, max_err_float, stream=None) -> float:
    out = ctypes.c_double()
    check(lib().hsx_max_err(ctypes.byref(g), dptr(u), dptr(v), rng(r), float(max_err_float),
                            ctypes.byref(out), stream))
    return out.value


def min_max(g: HsxGrid, u, r, stream=None):
    out = (ctypes.c_double * 2)()
    check(lib().hsx_min_max(ctypes.byref(g), dptr(u), rng(r), out, stream))
    return out[0], out[1]


def sync(stream=None):
    check(lib().hsx_stream_sync(stream))


# ---- modeling / adjoint ------------------------------------------------------------
def ac2_step_to(g: HsxGrid, out, prn, prc, coef, stream=None):
    check(lib().hsx_ac2_step_to(ctypes.byref(g), dptr(out), dptr(prn), dptr(prc), dptr(coef), stream))


def ac2_step_fs(g: HsxGrid, prn, prc, coef, fold, face_mask=ALL_FACES, stream=None):
    """Free surface folded into the step (hsx_ac2_step_fs). Raises HsxError with
    rc HSX_ERR_UNSUPPORTED when the grid does not take the folded kernel."""
    check(lib().hsx_ac2_step_fs(ctypes.byref(g), dptr(prn), dptr(prc), dptr(coef), face_mask, fold, stream))


def inject(g: HsxGrid, u, n, idx, val, w=None, scale=1.0, mode=HSX_INJECT_ADD, stream=None):
    check(lib().hsx_inject(ctypes.byref(g), dptr(u), int(n), dptr(idx), dptr(val), dptr(w), float(scale), mode,
                           stream))


def gather(g: HsxGrid, u, n, idx, out, stream=None):
    check(lib().hsx_gather(ctypes.byref(g), dptr(u), int(n), dptr(idx), dptr(out), stream))


def imaging(g: HsxGrid, grad, mu, u_next, u, u_prev, r, scale, src_idx=-1, src_val=0.0, stream=None):
    check(lib().hsx_imaging(ctypes.byref(g), dptr(grad), dptr(mu), dptr(u_next), dptr(u), dptr(u_prev), rng(r),
                            float(scale), int(src_idx), float(src_val), stream))


def ckpt_bytes(n: int) -> int:
    return int(lib().hsx_ckpt_bytes(int(n)))


def ckpt_encode(u, u_prev, n, packed, fmt, stream=None):
    check(lib().hsx_ckpt_encode(dptr(u), dptr(u_prev), int(n), dptr(packed), fmt, stream))


def ckpt_decode(packed, n, fmt, base, out, stream=None):
    check(lib().hsx_ckpt_decode(dptr(packed), int(n), fmt, dptr(base), dptr(out), stream))


def check_indices(idx, npoint: int, what: str = "index"):
    """Raise ValueError unless every flat grid index lies in [0, npoint)."""
    a = np.asarray(idx, dtype=np.int64).reshape(-1)
    if a.size and (a.min() < 0 or a.max() >= npoint):
        bad = a[(a < 0) | (a >= npoint)][0]
        raise ValueError(f"{what} {int(bad)} outside the grid's {npoint} points")


def make_shot(nt, dt, src_idx, wavelet, rec_idx, npoint: Optional[int] = None, host_src=None,
              host_rec=None) -> HsxShot:
    """Shot descriptor over DEVICE arrays (DeviceArray / tensors / addresses);
    the caller keeps them alive.  With npoint, the host copies of the indices
    (host_src / host_rec, or the tensors themselves) are checked first."""
    if npoint is not None:
        for what, dev, host in (("source index", src_idx, host_src), ("receiver index", rec_idx, host_rec)):
            if host is None and isinstance(dev, tuple) and torch is not None and isinstance(dev[0], torch.Tensor):
                host = dev[0].detach().cpu().numpy()
            if host is not None:
                check_indices(host, npoint, what)
    s = HsxShot()
    s.nt, s.dt = int(nt), float(dt)
    s.nsrc = int(src_idx[1]) if isinstance(src_idx, tuple) else 0
    s.src_idx = dptr(src_idx[0]).value if isinstance(src_idx, tuple) else None
    s.wavelet = dptr(wavelet).value if wavelet is not None else None
    s.nrec = int(rec_idx[1]) if isinstance(rec_idx, tuple) else 0
    s.rec_idx = dptr(rec_idx[0]).value if isinstance(rec_idx, tuple) else None
    return s


class Batch:
    """hsx_batch over lists of descriptors, coef buffers, shots and trace buffers
    (the caller keeps the buffers alive)."""

    def __init__(self, grids, coefs, shots, traces):
        n = len(grids)
        self._grids = (HsxGrid * n)(*grids)
        self._coef = (ctypes.c_void_p * n)(*[dptr(c).value for c in coefs])
        self._shots = (HsxShot * n)(*shots)
        self._traces = (ctypes.c_void_p * n)(*[dptr(t).value for t in traces])
        h = ctypes.c_void_p()
        check(lib().hsx_batch_create(n, self._grids, self._coef, ctypes.cast(self._shots, ctypes.c_void_p),
                                     self._traces, ctypes.byref(h)))
        self.h = h

    def run(self, use_graph=True, stream=None):
        check(lib().hsx_batch_run(self.h, 1 if use_graph else 0, stream))

    def set_shape(self, lanes, rows, chunk, ring):
        """hsx_batch_set_shape (chunk -1: one persistent launch per step)"""
        check(lib().hsx_batch_set_shape(self.h, int(lanes), int(rows), int(chunk), int(ring)))

    def autotune(self, reps=3):
        """hsx_batch_autotune: ((lanes, rows, chunk, ring), ms per step); chunk -1
        is the persistent launch"""
        shape = (ctypes.c_int * 4)()
        ms = ctypes.c_float(0.0)
        check(lib().hsx_batch_autotune(self.h, int(reps), shape, ctypes.byref(ms)))
        return tuple(shape), ms.value

    def close(self):
        if self.h is not None and self.h.value:
            lib().hsx_batch_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def forward(g: HsxGrid, coef, shot: HsxShot, traces, stream=None):
    check(lib().hsx_forward(ctypes.byref(g), dptr(coef), ctypes.byref(shot), dptr(traces), stream))


def _opts(ckpt_every, ckpt_format, residual, autotune=False):
    o = HsxAdjointOpts()
    o.ckpt_every, o.ckpt_format = int(ckpt_every), int(ckpt_format)
    o.flags = (HSX_ADJ_RESIDUAL if residual else 0) | (HSX_ADJ_AUTOTUNE if autotune else 0)
    return o


def gradient(g: HsxGrid, coef, shot: HsxShot, d_obs, grad, ckpt_every=0, ckpt_format=HSX_CKPT_FP32,
             stream=None, residual=False) -> float:
    """hsx_gradient; residual=True: d_obs is the residual (HSX_ADJ_RESIDUAL)."""
    o = _opts(ckpt_every, ckpt_format, residual)
    mis = ctypes.c_double(0.0)
    check(lib().hsx_gradient(ctypes.byref(g), dptr(coef), ctypes.byref(shot), dptr(d_obs), ctypes.byref(o),
                             dptr(grad), ctypes.byref(mis), stream))
    return mis.value


class AdjointPlan:
    """hsx_adjoint_plan: the gradient's work buffers allocated once for one
    grid layout, nt and receiver count."""

    def __init__(self, g: HsxGrid, nt, nrec, ckpt_every=0, ckpt_format=HSX_CKPT_FP32, residual=False,
                 autotune=False):
        self._o = _opts(ckpt_every, ckpt_format, residual, autotune)
        h = ctypes.c_void_p()
        check(lib().hsx_adjoint_plan_create(ctypes.byref(g), int(nt), int(nrec), ctypes.byref(self._o),
                                            ctypes.byref(h)))
        self.h = h

    def gradient(self, g: HsxGrid, coef, shot: HsxShot, d_obs, grad, stream=None) -> float:
        mis = ctypes.c_double(0.0)
        check(lib().hsx_gradient_plan(self.h, ctypes.byref(g), dptr(coef), ctypes.byref(shot), dptr(d_obs),
                                      dptr(grad), ctypes.byref(mis), stream))
        return mis.value

    def close(self):
        if self.h is not None and self.h.value:
            lib().hsx_adjoint_plan_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Event:
    def __init__(self):
        p = ctypes.c_void_p()
        check(lib().hsx_event_create(ctypes.byref(p)))
        self.ptr = p

    def record(self, stream=None):
        check(lib().hsx_event_record(self.ptr, stream))

    def elapsed_ms(self, later: "Event") -> float:
        ms = ctypes.c_float()
        check(lib().hsx_event_elapsed_ms(self.ptr, later.ptr, ctypes.byref(ms)))
        return ms.value

    def __del__(self):
        try:
            lib().hsx_event_destroy(self.ptr)
        except Exception:
            pass
