"""hpcscan_amd: MI355X-native implementation of hpcscan's high-order FD stencil path.

* ``hpcscan_amd/lib/libhpcscan_mi355x.so`` -- gfx950 kernels behind the C ABI
  declared in ``include/hpcscan_mi355x.h``.
* ``hpcscan_amd/bin/hpcscan[_fp64|_fp32c64]`` -- the C++ host harness with the
  reference's command line (``-testCase FD_D2|Propa -testMode MI355X``).
* :mod:`hpcscan_amd.capi` -- ctypes binding of the C ABI;
  :mod:`hpcscan_amd.layout` -- the Grid index layout for Python callers.
"""
import os

HERE = os.path.dirname(os.path.abspath(__file__))
BIN_DIR = os.path.join(HERE, "bin")
LIB_DIR = os.path.join(HERE, "lib")


def binary(precision: str = "fp32") -> str:
    """Path of the host harness: 'fp32', 'fp64' or 'fp32c64' (FP32 storage, FP64 compute)."""
    name = {"fp32": "hpcscan", "fp64": "hpcscan_fp64", "fp32c64": "hpcscan_fp32c64"}[precision]
    return os.path.join(BIN_DIR, name)
