"""Host logic that needs no GPU: the CPU build of the harness (oracle backends),
the product binaries' command line, and the C ABI library's exported symbols."""
import json
import os
import re
import subprocess

import pytest

from conftest import ROOT, report_value, run_bin

GOLDEN = os.path.join(ROOT, "tests", "golden")
CPU = os.path.join(ROOT, "oracle", "_build", "hpcscan_cpu")
CPU64 = os.path.join(ROOT, "oracle", "_build", "hpcscan_cpu_fp64")
HEADER = os.path.join(ROOT, "include", "hpcscan_mi355x.h")


def test_cpu_harness_reproduces_published_run(oracle, tmp_path):
    """misc/fileForReadme/runPropaTestCase.txt: -testCase Propa -testMode CacheBlk."""
    ref = json.load(open(os.path.join(GOLDEN, "propa_default_61.json")))
    rc, out = run_bin(CPU, ["-testCase", "Propa", "-testMode", "CacheBlk", "-ntry", "2"], tmp_path)
    assert rc == 0
    for k, v in ref.items():
        if k.startswith("_"):
            continue
        assert report_value(out, k) == v, (k, report_value(out, k), v)
    assert " Test PASSED" in out and "FAILED" not in out


def test_cpu_harness_perf_log_columns(oracle, tmp_path):
    run_bin(CPU, ["-testCase", "Propa", "-testMode", "Baseline", "-ntry", "1", "-nt", "5"], tmp_path)
    line = open(tmp_path / "hpcscan.perf.Propa.log").read().split()
    assert line[1:4] == ["Propa", "Baseline", "Ac2Standard"]
    assert len(line) == 4 + 9 + 14          # string columns, cols 1-9, cols 10-23
    assert line[4 + 8] == "8"               # col 9 fdOrder
    assert int(line[4 + 15]) == 5           # col 16 nt


@pytest.mark.parametrize("mode", ["Baseline", "CacheBlk"])
def test_cpu_validation_subset_zero_failed(oracle, tmp_path, mode):
    """Stencil part of script/testDriver.sh on the CPU oracle: 0 FAILED."""
    failed = passed = 0
    for tc in ("FD_D2", "Propa"):
        for dim in ("1", "2", "3"):
            for order in ("2", "6", "8", "16"):
                for extra in ([], ["-propagator", "Ac2SplitComp"]) if tc == "Propa" else ([],):
                    rc, out = run_bin(CPU, ["-testCase", tc, "-testMode", mode, "-ntry", "1", "-dim", dim,
                                            "-fdOrder", order] + extra, tmp_path)
                    assert rc == 0
                    passed += out.count("Test PASSED")
                    failed += out.count("Test FAILED")
    assert failed == 0 and passed > 0


@pytest.mark.parametrize("dim", ["1", "2", "3"])
def test_cpu_grid_test_case(oracle, tmp_path, dim):
    """-testCase Grid (src/testCase_Grid.cpp): nine operations, each checked
    against its closed-form answer; 18 perf-log columns after the common ones."""
    rc, out = run_bin(CPU, ["-testCase", "Grid", "-testMode", "CacheBlk", "-ntry", "2", "-dim", dim, "-n1", "37",
                            "-n2", "29", "-n3", "21", "-fdOrder", "8"], tmp_path)
    assert rc == 0
    cases = re.findall(r"\* Case\s+(\S+)", out)
    assert cases == ["Grid" + c for c in ("Fill", "MaxErr", "L1Err", "GetSumAbs", "GetSumAbsDiff", "GetMax",
                                          "GetMin", "UpdatePressure", "ApplyBoundaryCondition")]
    assert out.count("Test PASSED") == 9 and "FAILED" not in out
    line = open(tmp_path / "hpcscan.perf.Grid.log").read().split()
    assert len(line) == 4 + 9 + 18


@pytest.mark.parametrize("dim", ["1", "2", "3"])
@pytest.mark.parametrize("order", ["2", "8", "16"])
def test_cpu_fd_d1_test_case(oracle, tmp_path, dim, order):
    """-testCase FD_D1 (src/testCase_FD_D1.cpp): one staggered first derivative per
    axis against the analytic cosine half a step back, L1 < 1e-3; 25 perf-log
    columns after the common ones."""
    rc, out = run_bin(CPU, ["-testCase", "FD_D1", "-testMode", "CacheBlk", "-ntry", "2", "-dim", dim,
                            "-fdOrder", order], tmp_path)
    assert rc == 0
    cases = re.findall(r"\* Case\s+(\S+)", out)
    assert cases == ["FD_D1Axis%d" % (k + 1) for k in range(int(dim))]
    assert out.count("Test PASSED") == int(dim) and "FAILED" not in out
    assert report_value(out, "#Point in stencil") == order
    line = open(tmp_path / "hpcscan.perf.FD_D1.log").read().split()
    assert len(line) == 4 + 9 + 25


def write_vp_model(path, n, vmin=1500.0, grad=40.0, fp64=False):
    """velocity increasing with depth (n1 first) in the storage precision (Myfloat),
    the inner grid of the command line with n1 fastest, as Grid::read expects
    (src/grid.cpp:1587-1643)"""
    import numpy as np

    t = np.float64 if fp64 else np.float32
    n = list(n) + [1] * (3 - len(n))
    v = (vmin + grad * np.arange(n[0])).astype(t)
    np.broadcast_to(v[None, None, :], (n[2], n[1], n[0])).astype(t).tofile(path)


MODELING_ARGS = ["-testCase", "Modeling", "-fdOrder", "8", "-d1", "10", "-d2", "10", "-d3", "10", "-freqMax", "20",
                 "-nt", "120", "-ntry", "1"]


@pytest.mark.parametrize("dim,n,nrec", [("1", (60,), 1), ("2", (40, 50), 50), ("3", (40, 50, 60), 300)])
def test_cpu_modeling_test_case(oracle, tmp_path, dim, n, nrec):
    """-testCase Modeling (src/testCase_Modeling.cpp): Vp model read from disk,
    Ricker source set one point below the top, receivers on that plane; the
    trace file holds nt samples per receiver, the .info file nt, nRec, 1."""
    import numpy as np

    write_vp_model(tmp_path / "vp.bin", n)
    size = []
    for k, a in zip(("-n1", "-n2", "-n3"), n):
        size += [k, str(a)]
    rc, out = run_bin(CPU, MODELING_ARGS + ["-testMode", "CacheBlk", "-dim", dim, "-modelVpFile",
                                            str(tmp_path / "vp.bin")] + size, tmp_path)
    assert rc == 0, out[-2000:]
    assert report_value(out, "Number of receivers") == str(nrec)
    assert open(tmp_path / "ModelingForwardPrn.trace.info").read().split() == ["120", str(nrec), "1"]
    t = np.fromfile(tmp_path / "ModelingForwardPrn.trace.bin", np.float32).reshape(nrec, 120)
    assert np.isfinite(t).all() and np.abs(t).max() > 0.1
    line = open(tmp_path / "hpcscan.perf.Modeling.log").read().split()
    assert len(line) == 4 + 9 + 22


def test_cpu_modeling_requires_a_model(oracle, tmp_path):
    rc, out = run_bin(CPU, ["-testCase", "Modeling", "-testMode", "CacheBlk", "-nt", "5"], tmp_path)
    assert "modelVpFile has not been specified" in out


def test_split_and_standard_print_same_error(oracle, tmp_path):
    outs = []
    for prop in ("Ac2Standard", "Ac2SplitComp"):
        rc, out = run_bin(CPU, ["-testCase", "Propa", "-testMode", "CacheBlk", "-ntry", "1", "-n1", "40", "-n2",
                                "40", "-n3", "40", "-nt", "60", "-propagator", prop], tmp_path)
        outs.append(report_value(out, "errTestCase2"))
    assert outs[0] == outs[1]


def test_cpu_harness_fp64(oracle, tmp_path):
    rc, out = run_bin(CPU64, ["-testCase", "Propa", "-testMode", "Baseline", "-ntry", "1"], tmp_path)
    assert rc == 0 and "FP64 [8 Bytes]" in out and " Test PASSED" in out


def _product_bins():
    from hpcscan_amd import binary

    return [binary(p) for p in ("fp32", "fp64", "fp32c64")]


def _have_gpu():
    try:
        from hpcscan_amd import capi

        return capi.device_count() > 0
    except Exception:
        return False


@pytest.mark.parametrize("path", _product_bins())
def test_product_cli(tmp_path, path):
    if not os.path.exists(path):
        pytest.fail(f"{path} is not built: run __graft_entry__.build()")
    rc, out = run_bin(path, ["-h"], tmp_path)
    assert rc == 0
    for flag in ("-testCase", "-testMode", "-fdOrder", "-propagator", "-hipKernel", "-gpuDevice", "-snapDt"):
        assert flag in out
    rc, out = run_bin(path, ["-v"], tmp_path)
    assert "Mode MI355X\tENABLED" in out and "Mode Baseline" not in out   # no CPU backend in the product
    rc, out = run_bin(path, ["-testCase", "Propa", "-fdOrder", "5"], tmp_path)
    assert rc != 0 and "FD space order should be" in out
    rc, out = run_bin(path, ["-testCase", "Propa", "-bogus"], tmp_path)
    assert rc != 0 and "Unknown argument" in out
    rc, out = run_bin(path, ["-testCase", "Propa", "-testMode", "Baseline"], tmp_path)
    assert "not supported or invalid gridMode" in out


def test_product_fails_loudly_without_gpu(tmp_path):
    from hpcscan_amd import binary

    if _have_gpu():
        pytest.skip("a GPU is visible: the no-GPU error path cannot be exercised here")
    rc, out = run_bin(binary(), ["-testCase", "Propa", "-testMode", "MI355X"], tmp_path)
    assert "hsx_init failed" in out and "F A T A L" in out
    assert "Test PASSED" not in out


def test_c_abi_exports_every_declared_symbol():
    """The library loads (after torch, sharing its HIP runtime) and exports every
    entry point of include/hpcscan_mi355x.h; no compute call is made."""
    from hpcscan_amd import capi

    L = capi.lib()
    text = open(HEADER).read()
    declared = set(re.findall(r"\b(hsx_[a-z0-9_]+)\s*\(", text))
    assert len(declared) >= 30
    for name in declared:
        assert hasattr(L, name), name
    assert declared == set(capi.SIGNATURES), declared ^ set(capi.SIGNATURES)
    assert L.hsx_abi_version() == 1
    # symbols really exported by the shared object, checked with nm
    nm = subprocess.run(["nm", "-D", "--defined-only", capi.LIB_PATH], capture_output=True, text=True).stdout
    for name in declared:
        assert re.search(rf"\bT {name}$", nm, re.M), name


def test_c_abi_argument_errors_without_gpu():
    """Descriptor validation happens on the host before any device work."""
    from hpcscan_amd import capi
    from hpcscan_amd.layout import make_layout

    if _have_gpu():
        pytest.skip("GPU visible: covered by the GPU tests")
    g = capi.hsx_grid(make_layout(3, (10, 10, 10), 8))
    rc = capi.lib().hsx_ac2_step(None, None, None, None, None)
    assert rc != 0
    assert capi.last_error()


@pytest.mark.parametrize("argv,env", [(["--gpus", "2"], {}), ([], {"WORLD_SIZE": "2"})])
def test_bench_refuses_more_than_one_gpu(argv, env):
    """bench.py measures one GPU: more ranks or --gpus > 1 end with exit code 2
    and a message, before any device is touched"""
    import sys

    e = dict(os.environ, **env)
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py")] + argv, capture_output=True, text=True,
                       timeout=120, env=e)
    assert r.returncode == 2
    assert "one MI355X" in r.stderr and r.stdout == ""
