"""The hpcscan command line on -testMode MI355X (run with -m gpu).

The product binaries (hpcscan_amd/bin) run the reference's test cases on the
GPU; their printed results are compared with the reference's published
output and with the oracle build of the same harness on the CPU
(-testMode CacheBlk), line by line for the error values.
"""
import json
import os

import pytest

from conftest import ROOT, report_value, run_bin
from hpcscan_amd import binary

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(ROOT, "tests", "golden")
CPU_BIN = {"fp32": os.path.join(ROOT, "oracle", "_build", "hpcscan_cpu"),
           "fp64": os.path.join(ROOT, "oracle", "_build", "hpcscan_cpu_fp64")}


def _errors(out):
    """every printed error / PASS / FAIL line, in order"""
    keep = ("errTestCase2", "Test PASSED", ">> Test FAILED", "ntCheck", "CFL", "Stable time step",
            "No. time steps", "Time step (s)")
    return [l.strip() for l in out.splitlines() if l.strip().startswith(keep)]


def test_default_propa_matches_published(tmp_path):
    ref = json.load(open(os.path.join(GOLDEN, "propa_default_61.json")))
    rc, out = run_bin(binary(), ["-testCase", "Propa", "-testMode", "MI355X", "-ntry", "2"], tmp_path)
    print(out[-3000:])
    assert rc == 0
    for k in ("errTestCase2", "CFL", "Stable time step (s)", "Total n1", "Grid Inner Glob. (Pts)", "ntCheck",
              "#Flop per point", "#Point in stencil", "Max. frequency (Hz)"):
        assert report_value(out, k) == ref[k], k
    assert ">> Test FAILED" not in out
    assert "Mode MI355X\tENABLED" in out


@pytest.mark.parametrize("prec", ["fp32", "fp64"])
@pytest.mark.parametrize("args", [
    ["-testCase", "Propa", "-dim", "1", "-fdOrder", "4"],
    ["-testCase", "Propa", "-dim", "2", "-fdOrder", "12"],
    ["-testCase", "Propa", "-dim", "3", "-fdOrder", "16"],
    ["-testCase", "Propa", "-dim", "3", "-fdOrder", "6", "-propagator", "Ac2SplitComp"],
    ["-testCase", "Propa", "-dim", "3", "-nt", "50", "-dt", "0.001", "-param1", "2", "-param3", "2",
     "-snapInc", "20"],
    ["-testCase", "Propa", "-n1", "100", "-n2", "100", "-n3", "100", "-fdOrder", "8", "-nt", "100",
     "-ratioCFL", "0.8", "-param1", "1", "-param2", "1", "-param3", "1"],
    ["-testCase", "FD_D2", "-dim", "3", "-fdOrder", "8"],
    ["-testCase", "FD_D2", "-dim", "3", "-fdOrder", "16", "-n1", "27", "-n2", "36", "-n3", "41"],
    ["-testCase", "FD_D2", "-dim", "2", "-fdOrder", "2", "-n1AddPad", "5", "-n2AddPad", "5"],
    ["-testCase", "FD_D2", "-n1Offset", "1", "-n2Offset", "2", "-n1MulPad", "5"],
    ["-testCase", "FD_D1", "-dim", "3", "-fdOrder", "8"],
    ["-testCase", "FD_D1", "-dim", "3", "-fdOrder", "16", "-n1", "68", "-n2", "36", "-n3", "41"],
    ["-testCase", "FD_D1", "-dim", "2", "-fdOrder", "2"],
    ["-testCase", "FD_D1", "-dim", "1", "-fdOrder", "12"],
])
def test_same_report_as_cpu_reference(tmp_path, prec, args):
    """Error values printed by the MI355X run equal those of the CPU run."""
    print(" ".join(args), prec, flush=True)
    rc1, gpu = run_bin(binary(prec), args + ["-testMode", "MI355X", "-ntry", "1"], tmp_path)
    rc2, cpu = run_bin(CPU_BIN[prec], args + ["-testMode", "CacheBlk", "-ntry", "1"], tmp_path)
    assert rc1 == 0 and rc2 == 0
    assert _errors(gpu) == _errors(cpu)
    assert ">> Test FAILED" not in gpu


@pytest.mark.parametrize("prec", ["fp32", "fp64"])
@pytest.mark.parametrize("args", [
    ["-testCase", "Propa", "-dim", "3", "-fdOrder", "8"],
    ["-testCase", "Propa", "-dim", "3", "-fdOrder", "4", "-n1", "72", "-n2", "40", "-n3", "36", "-nt", "40"],
    ["-testCase", "Propa", "-dim", "3", "-fdOrder", "16", "-nt", "30"],
    ["-testCase", "Propa", "-dim", "3", "-fdOrder", "8", "-n1", "100", "-n2", "100", "-n3", "100", "-nt", "60",
     "-param1", "1", "-param2", "1", "-param3", "1", "-snapInc", "7"],
    ["-testCase", "Propa", "-dim", "2", "-fdOrder", "8"],          # 2D: the free surface stays apart
])
def test_folded_free_surface_same_report(tmp_path, prec, args):
    """-hipFoldBC (free surface folded into the fused step, its boundary-plane
    zeros owed until the grid is used otherwise) prints the CPU run's errors."""
    print(" ".join(args), prec, flush=True)
    rc1, gpu = run_bin(binary(prec), args + ["-testMode", "MI355X", "-ntry", "2", "-hipFoldBC"], tmp_path)
    rc2, cpu = run_bin(CPU_BIN[prec], args + ["-testMode", "CacheBlk", "-ntry", "1"], tmp_path)
    assert rc1 == 0 and rc2 == 0
    assert _errors(gpu) == _errors(cpu)
    assert ">> Test FAILED" not in gpu and "F A T A L" not in gpu


@pytest.mark.parametrize("prec", ["fp32", "fp64"])
@pytest.mark.parametrize("dim,n,extra", [
    ("3", (40, 50, 60), []),
    ("3", (72, 40, 36), ["-fdOrder", "16"]),
    ("3", (72, 40, 36), ["-hipFoldBC"]),
    ("3", (64, 48, 40), ["-propagator", "Ac2SplitComp"]),
    ("2", (40, 50), []),
    ("1", (60,), []),
])
def test_modeling_traces_equal_cpu(tmp_path, prec, dim, n, extra):
    """-testCase Modeling on the MI355X backend (device source term, receivers
    and traces): the trace file equals the CPU run's byte for byte."""
    from test_harness_cpu import MODELING_ARGS, write_vp_model

    print("modeling", dim, n, extra, prec, flush=True)
    write_vp_model(tmp_path / "vp.bin", n, fp64=prec == "fp64")
    size = []
    for k, a in zip(("-n1", "-n2", "-n3"), n):
        size += [k, str(a)]
    args = MODELING_ARGS + ["-dim", dim, "-modelVpFile", str(tmp_path / "vp.bin")] + size
    gdir, cdir = tmp_path / "gpu", tmp_path / "cpu"
    gdir.mkdir()
    cdir.mkdir()
    gpu_extra = [x for x in extra if x == "-hipFoldBC"]
    both = [x for x in extra if x != "-hipFoldBC"]
    rc1, gpu = run_bin(binary(prec), args + both + gpu_extra + ["-testMode", "MI355X"], gdir)
    rc2, cpu = run_bin(CPU_BIN[prec], args + both + ["-testMode", "CacheBlk"], cdir)
    assert rc1 == 0 and rc2 == 0, gpu[-2000:]
    g = open(gdir / "ModelingForwardPrn.trace.bin", "rb").read()
    c = open(cdir / "ModelingForwardPrn.trace.bin", "rb").read()
    assert len(g) == len(c) > 0
    assert g == c
    assert "F A T A L" not in gpu


@pytest.mark.parametrize("prec", ["fp32", "fp64"])
@pytest.mark.parametrize("dim", ["1", "2", "3"])
def test_grid_test_case_on_gpu(tmp_path, prec, dim):
    """-testCase Grid on the MI355X backend: the nine operations pass their
    closed-form checks, as on the CPU oracle build"""
    args = ["-testCase", "Grid", "-dim", dim, "-n1", "64", "-n2", "48", "-n3", "40", "-fdOrder", "8", "-ntry", "2"]
    rc, gpu = run_bin(binary(prec), args + ["-testMode", "MI355X"], tmp_path)
    rc2, cpu = run_bin(CPU_BIN[prec], args + ["-testMode", "CacheBlk"], tmp_path)
    assert rc == 0 and rc2 == 0
    assert gpu.count("Test PASSED") == 9 and "FAILED" not in gpu, gpu[-3000:]
    assert cpu.count("Test PASSED") == 9


def test_validation_matrix_zero_failed(tmp_path):
    """The stencil part of the reference's validation driver (script/testDriver.sh:79-141,
    233-321) on -testMode MI355X, one rank: 0 FAILED."""
    runs = []
    for tc in ("FD_D1", "FD_D2", "Propa"):
        for dim in ("1", "2", "3"):
            for order in ("2", "4", "6", "8", "10", "12", "14", "16"):
                runs.append(["-testCase", tc, "-dim", dim, "-fdOrder", order])
                if tc == "Propa":
                    runs.append(["-testCase", tc, "-dim", dim, "-fdOrder", order, "-propagator", "Ac2SplitComp"])
    for order in ("2", "8", "16"):
        runs.append(["-testCase", "FD_D2", "-fdOrder", order, "-dim", "3", "-autoPad"])
    runs.append(["-testCase", "Propa", "-dim", "1", "-nt", "50", "-dt", "0.001", "-fdOrder", "8", "-n1", "100",
                 "-param1", "2", "-snapInc", "20"])
    runs.append(["-testCase", "Propa", "-dim", "2", "-nt", "50", "-dt", "0.001", "-fdOrder", "8", "-n1", "100",
                 "-n2", "100", "-param2", "2", "-param1", "2", "-snapInc", "20"])
    passed = failed = 0
    for a in runs:
        rc, out = run_bin(binary(), a + ["-testMode", "MI355X", "-ntry", "1"], tmp_path)
        assert rc == 0, a
        passed += out.count("Test PASSED")
        failed += out.count("Test FAILED")
        assert "F A T A L" not in out, (a, out[-2000:])
    print(f"validation matrix: {len(runs)} runs, {passed} PASSED, {failed} FAILED")
    assert failed == 0 and passed >= len(runs)


@pytest.mark.slow
def test_published_1000_cube_error(tmp_path):
    """Propa 1000^3 O8 -ratioCFL 0.5, modes 130, tmax 0.1, snapDt 0.01: the
    reference's CPU run printed errTestCase 0.0340336 (nt 451)."""
    args = ["-testCase", "Propa", "-testMode", "MI355X", "-ntry", "1", "-n1", "1000", "-n2", "1000", "-n3", "1000",
            "-fdOrder", "8", "-ratioCFL", "0.5", "-tmax", "0.1", "-snapDt", "0.01", "-param1", "130",
            "-param2", "130", "-param3", "130"]
    rc, out = run_bin(binary(), args, tmp_path, timeout=900)
    print(out[-2500:])
    assert rc == 0
    assert report_value(out, "No. time steps") == "451"
    err = float(report_value(out, "errTestCase2"))
    assert abs(err / 0.0340336 - 1) < 5e-5


@pytest.mark.parametrize("prec", ["fp32", "fp64"])
def test_autotune_flag_same_report(tmp_path, prec):
    """-hipAutotune: the backend tunes the tile shape of the grid layout once
    (printed), and the errors equal the CPU run's"""
    args = ["-testCase", "Propa", "-dim", "3", "-n1", "100", "-n2", "96", "-n3", "90", "-fdOrder", "8", "-nt", "30"]
    rc1, gpu = run_bin(binary(prec), args + ["-testMode", "MI355X", "-ntry", "1", "-hipAutotune"], tmp_path)
    rc2, cpu = run_bin(CPU_BIN[prec], args + ["-testMode", "CacheBlk", "-ntry", "1"], tmp_path)
    assert rc1 == 0 and rc2 == 0
    assert report_value(gpu, "MI355X autotuned tile") is not None, gpu[-3000:]
    print(report_value(gpu, "MI355X autotuned tile"))
    assert _errors(gpu) == _errors(cpu)


@pytest.mark.parametrize("prec", ["fp32", "fp64"])
@pytest.mark.parametrize("args", [
    ["-testCase", "Propa", "-dim", "3", "-n1", "40", "-n2", "36", "-n3", "32", "-fdOrder", "8", "-nt", "12",
     "-snapInc", "4"],
    ["-testCase", "Propa", "-dim", "2", "-n1", "44", "-n2", "38", "-fdOrder", "4", "-nt", "10",
     "-propagator", "Ac2SplitComp"],
    ["-testCase", "FD_D2", "-dim", "3", "-n1", "30", "-n2", "28", "-n3", "26", "-fdOrder", "12"],
])
def test_write_grid_files_equal_cpu(tmp_path, prec, args):
    """-writeGrid (Grid::write, src/grid.cpp:1245-1310): every .grid.info /
    .grid.bin file the MI355X run writes equals the CPU run's byte for byte"""
    gdir, cdir = tmp_path / "gpu", tmp_path / "cpu"
    gdir.mkdir()
    cdir.mkdir()
    rc1, gpu = run_bin(binary(prec), args + ["-testMode", "MI355X", "-ntry", "1", "-writeGrid"], gdir)
    rc2, cpu = run_bin(CPU_BIN[prec], args + ["-testMode", "CacheBlk", "-ntry", "1", "-writeGrid"], cdir)
    assert rc1 == 0 and rc2 == 0, gpu[-2000:]
    gfiles = sorted(f for f in os.listdir(gdir) if ".grid." in f)
    cfiles = sorted(f for f in os.listdir(cdir) if ".grid." in f)
    print(prec, " ".join(args), gfiles, flush=True)
    assert gfiles == cfiles and len(gfiles) >= 2
    assert any(f.endswith(".grid.bin") for f in gfiles) and any(f.endswith(".grid.info") for f in gfiles)
    for f in gfiles:
        a, b = open(gdir / f, "rb").read(), open(cdir / f, "rb").read()
        assert len(a) == len(b) > 0, f
        assert a == b, f


@pytest.mark.parametrize("prec", ["fp32", "fp64"])
def test_write_global_snapshots_equal_cpu(tmp_path, prec):
    """-testCase Modeling -writeGrid: the snapshots written by writeGlobal
    (src/grid.cpp:1358-1441) equal the CPU run's byte for byte"""
    from test_harness_cpu import MODELING_ARGS, write_vp_model

    n = (48, 40, 36)
    write_vp_model(tmp_path / "vp.bin", n, fp64=prec == "fp64")
    args = MODELING_ARGS + ["-dim", "3", "-modelVpFile", str(tmp_path / "vp.bin"), "-n1", "48", "-n2", "40",
                            "-n3", "36", "-writeGrid"]
    gdir, cdir = tmp_path / "gpu", tmp_path / "cpu"
    gdir.mkdir()
    cdir.mkdir()
    rc1, gpu = run_bin(binary(prec), args + ["-testMode", "MI355X"], gdir)
    rc2, cpu = run_bin(CPU_BIN[prec], args + ["-testMode", "CacheBlk"], cdir)
    assert rc1 == 0 and rc2 == 0, gpu[-2000:]
    gfiles = sorted(f for f in os.listdir(gdir) if ".grid." in f)
    cfiles = sorted(f for f in os.listdir(cdir) if ".grid." in f)
    print(prec, gfiles, flush=True)
    assert gfiles == cfiles and any(f.endswith(".grid.bin") for f in gfiles)
    for f in gfiles:
        a, b = open(gdir / f, "rb").read(), open(cdir / f, "rb").read()
        assert len(a) == len(b) > 0 and a == b, f


@pytest.mark.parametrize("args", [
    ["-testCase", "Propa", "-dim", "3", "-n1", "100", "-n2", "100", "-n3", "100", "-fdOrder", "16", "-nt", "60",
     "-param1", "2", "-param2", "2", "-param3", "2"],
    ["-testCase", "Propa", "-dim", "3", "-n1", "100", "-n2", "100", "-n3", "100", "-fdOrder", "12", "-nt", "60",
     "-propagator", "Ac2SplitComp", "-param1", "2", "-param2", "2", "-param3", "2"],
])
def test_contract_flag_close_to_cpu(tmp_path, args):
    """-hipContract (fused multiply-adds in the fused step and the Laplacian):
    the printed error stays within 3 % of the CPU run's and the test passes.
    (In FP32 the rounding of the stencil sums, whose terms cancel to (a h)^2 of
    their size, moves errTestCase by several percent; contraction rounds them
    differently: 1.0 % at order 16 here.  The wavefield itself stays within
    1e-5 relative L2 of the exact one, test_contracted_propagation_within_1e5.)"""
    rc1, gpu = run_bin(binary("fp32"), args + ["-testMode", "MI355X", "-ntry", "1", "-hipContract"], tmp_path)
    rc2, cpu = run_bin(CPU_BIN["fp32"], args + ["-testMode", "CacheBlk", "-ntry", "1"], tmp_path)
    assert rc1 == 0 and rc2 == 0
    eg, ec = float(report_value(gpu, "errTestCase2")), float(report_value(cpu, "errTestCase2"))
    print(" ".join(args), f"contracted {eg:.6g} CPU {ec:.6g}", flush=True)
    assert "Test PASSED" in gpu and abs(eg / ec - 1) < 3e-2
