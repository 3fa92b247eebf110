#!/usr/bin/env python3
"""Headline benchmark: Ac2Standard propagation, FP32, FD order 8, 1000^3, one MI355X.

One step is one time step of hpcscan's propagator (src/propagator_Ac2.cpp:366-393):
free-surface boundary on prc, then the fused Laplacian + leapfrog on prn, over
the 1000^3 inner grid (eigen mode 2,2,2 initial state, dt = CFL limit); the
free surface runs as launches of its own before each step (--fold: folded
into the fused kernel instead, slower on this device).  Grids are resident in
HBM before timing.  Prints one JSON line.

  value      GPoint/s = inner points x steps / wall time of the timed steps
  roofline   the fused step kernel: 16 algorithmic bytes per inner point
             (read prc, prn, coef; write prn) / its average HIP-event duration
  cpu_baseline  the oracle's CacheBlk restatement on this host's cores, on a
             time-bounded sample of the same workload
  analytical_error  errTestCase2 (error against the analytical eigenmode) of
             the MI355X harness on the CPU baseline's command line beside the
             CPU path's, and the timed wavefield's own L1 error at its last step
  batch      32 forward shots of uneven sizes in one batch (HIP graph)
  adjoint    the second half of the metric: seconds per shot of the misfit
             gradient (200^3 order 8, 1000 steps, 16-bit checkpoints) against
             the plain PyTorch full-storage version, with both gradients'
             error against the FP64 gradient (see adjoint_leg)
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Ac2 propagator GPoint/s & %HBM roofline; adjoint gradient s/shot vs plain"
HBM_PEAK_GBS = 8000.0
PI_REF = 3.1415926535897
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic_ac2_fp32_o8_1000.json")


def log(msg):
    sys.stderr.write(f"[bench] {msg}\n")
    sys.stderr.flush()


def propa_scalars(n, order):
    """stableDt / dt of Propagator_Ac2::initialize in float32 (ratioCFL 1)."""
    import numpy as np

    f32 = np.float32
    # sum |A_j| over the Myfloat2 (float) coefficients, getCFL (src/propagator_Ac2.cpp:410-439)
    table = {8: [-205 / 72, 8 / 5, -1 / 5, 8 / 315, -1 / 560], 4: [-5 / 2, 4 / 3, -1 / 12],
             16: [-1077749 / 352800, 16 / 9, -14 / 45, 112 / 1485, -7 / 396, 112 / 32175, -2 / 3861,
                  16 / 315315, -1 / 411840],
             12: [-5369 / 1800, 12 / 7, -15 / 56, 10 / 189, -1 / 112, 2 / 1925, -1 / 16632]}[order]
    s2 = abs(float(f32(table[0]))) + 2 * sum(abs(float(f32(a))) for a in table[1:])
    cfl = f32(math.sqrt(4.0 / (3 * s2)))
    d = f32(1.0 / (n - 1))
    stable = f32(f32(cfl * d) / f32(1.0))
    return float(stable)


def eigen_tables(p, modes):
    from hpcscan_amd.layout import unit_origin

    r = p.all_points()
    tabs = []
    for a in range(3):
        o = unit_origin(p, a)
        tabs.append([math.sin((o + i * p.d[a]) * (PI_REF * modes[a])) for i in range(r[2 * a], r[2 * a + 1] + 1)])
    return tabs


def gpu_leg(args):
    import numpy as np

    from hpcscan_amd import capi
    from hpcscan_amd.layout import make_layout

    capi.init(0)
    capi.set_kernel({"auto": capi.HSX_KERNEL_AUTO, "ref": capi.HSX_KERNEL_REF,
                     "tiled": capi.HSX_KERNEL_TILED}[args.kernel])
    capi.set_contract(args.contract)
    capi.set_tile(args.tile_lx, args.tile_chunk, args.tile_rows,
                  {"default": 0, "cached": capi.HSX_TILE_CACHED, "nt": capi.HSX_TILE_NT}[args.streams]
                  | (capi.HSX_TILE_NO_RING if args.no_ring else 0))
    info = capi.device_info(0)
    n, order = args.n, args.order
    p = make_layout(3, (n, n, n), order)
    g = capi.hsx_grid(p)
    npt = p.n_inner_points()
    log(f"device {info['name']} ({info['arch']}, {info['compute_units']} CUs); grid {p.n1}x{p.n2}x{p.n3}")

    dt = np.float32(propa_scalars(n, order))
    modes = (2, 2, 2)
    gamma = math.sqrt(12)
    tabs = eigen_tables(p, modes)
    r_all, r_in = p.all_points(), p.inner()
    nbytes = p.npoint * 4
    prn, prc, coef = (capi.DeviceArray(nbytes) for _ in range(3))
    full = [0, p.n1 - 1, 0, p.n2 - 1, 0, p.n3 - 1]
    for x in (prn, prc, coef):
        capi.fill_const(g, x, full, 0.0)
    capi.fill_const(g, coef, r_in, float(np.float32(dt * dt)))
    capi.fill_separable(g, prn, r_all, *tabs, math.sin(PI_REF * float(np.float32(-2.0 * dt)) * gamma))
    capi.fill_separable(g, prc, r_all, *tabs, math.sin(PI_REF * float(np.float32(-dt)) * gamma))
    capi.sync()
    tuned = None
    forced = (args.tile_lx, args.tile_rows, args.tile_chunk, args.streams, args.no_ring) != (0, 0, 0, "default", False)
    if args.autotune and not forced and args.kernel == "auto":
        # the tuner times its candidate shapes (the library's default first) on
        # scratch grids of this layout and keeps the fastest; setup, not timed
        shape, tms = capi.autotune(g, reps=5)
        tuned = {"lanes": shape[0], "rows": shape[1], "chunk": shape[2], "flags": shape[3], "ms": round(tms, 4)}
        log(f"autotuned tile {shape} ({tms:.3f} ms per step in the sweep)")

    # a step is the reference's free surface on prc then the fused update of prn;
    # folded (--fold): the first step applies the free surface to the initial
    # prc and writes the new field's halos, every later step reads the owed
    # boundary-plane zeros of prc / prn inside the fused kernel
    nstep = [0]
    folded = args.fold

    def step(a, b, ev=None):
        if folded:
            fold = capi.HSX_FOLD_HALO
            if nstep[0] == 0:
                capi.free_surface(g, b)
            else:
                fold |= capi.HSX_FOLD_PRC | capi.HSX_FOLD_PRN
            if ev:
                ev[0].record()
            capi.ac2_step_fs(g, a, b, coef, fold)
        else:
            capi.free_surface(g, b)
            if ev:
                ev[0].record()
            capi.ac2_step(g, a, b, coef)
        if ev:
            ev[1].record()
        nstep[0] += 1

    if args.calibrate_copy:
        # one plain copy of a whole grid (npoint*4 bytes read + written): the
        # reference kernel for scaling the FETCH_SIZE / WRITE_SIZE counters
        scratch = capi.DeviceArray(nbytes)
        capi.array_op(g, capi.HSX_ARRAY_COPY, scratch, prc)
        capi.sync()
        scratch.free()

    a, b = prn, prc
    for _ in range(args.warmup):
        step(a, b)
        a, b = b, a
    capi.sync()
    log(f"warmup {args.warmup} steps done")

    events = [(capi.Event(), capi.Event()) for _ in range(args.steps)]
    capi.sync()
    t0 = time.perf_counter()
    for k in range(args.steps):
        step(a, b, events[k])
        a, b = b, a
    capi.sync()
    t1 = time.perf_counter()
    kern_ms = [e0.elapsed_ms(e1) for e0, e1 in events]
    wall = t1 - t0
    log(f"timed {args.steps} steps in {wall:.3f} s")

    # sanity: the wavefield is finite and of the expected size
    s = capi.sum_abs(g, b, r_in)
    if not math.isfinite(s) or s <= 0:
        raise RuntimeError(f"wavefield is not finite after the run (sum|p| = {s})")
    # the timed run's own error against the analytical eigenmode: the newest
    # field is step it = warmup + steps - 1, time it * dt (in storage precision,
    # as src/testCase_Propa.cpp:94-106), relative L1 over the inner points
    it_last = args.warmup + args.steps - 1
    t_last = float(np.float32(np.float32(it_last) * dt))
    refg = capi.DeviceArray(nbytes)
    capi.fill_const(g, refg, full, 0.0)
    capi.fill_separable(g, refg, r_all, *tabs, math.sin(PI_REF * t_last * gamma))
    l1_last = capi.sum_abs_diff(g, b, refg, r_in) / capi.sum_abs(g, refg, r_in)
    refg.free()
    log(f"analytical error of the timed wavefield (step {it_last}, t = {t_last:.6g} s): relative L1 {l1_last:.6g}")

    # pointwise reference on the same grid: updatePressure (prn = 2 prc - prn +
    # coef * lapla, 4 reads + 1 write, no stencil, one thread per 16 bytes in
    # address order, nontemporal) -- what plain streaming of this layout
    # reaches on this device
    pw_ms = None
    if args.pointwise_ref:
        lapla = capi.DeviceArray(nbytes)
        capi.array_op(g, capi.HSX_ARRAY_COPY, lapla, coef)
        pev = [(capi.Event(), capi.Event()) for _ in range(6)]
        for e0, e1 in pev:
            e0.record()
            capi.update_pressure(g, a, b, coef, lapla)
            e1.record()
        capi.sync()
        pw_ms = sorted(e0.elapsed_ms(e1) for e0, e1 in pev[1:])[2]
        lapla.free()

    avg_kernel_s = sum(kern_ms) / len(kern_ms) / 1e3
    alg_bytes = 16 * npt
    achieved = alg_bytes / avg_kernel_s / 1e9
    traffic = None
    # the shape the timed launches used (forced, else tuned), as command-line
    # arguments for the child runs that count its HBM bytes
    if tuned is not None:
        lx, by, ch, fl = tuned["lanes"], tuned["rows"], tuned["chunk"], tuned["flags"]
        streams = "cached" if fl & capi.HSX_TILE_CACHED else "nt" if fl & capi.HSX_TILE_NT else "default"
        no_ring = bool(fl & capi.HSX_TILE_NO_RING)
    else:
        lx, by, ch, streams, no_ring = args.tile_lx, args.tile_rows, args.tile_chunk, args.streams, args.no_ring
    shape_args = ["--tile-lx", str(lx), "--tile-rows", str(by), "--tile-chunk", str(ch), "--streams", streams] + \
        (["--no-ring"] if no_ring else [])
    # the library's own shape at 1000^3 order 8 is 128 x 8 with the ring (the
    # committed counter summary is for it)
    default_shape = (lx, by, ch, streams, no_ring) in ((0, 0, 0, "default", False), (128, 8, 0, "default", False))
    if os.path.exists(TRAFFIC_FILE) and (n, order, args.kernel) == (1000, 8, "auto") and default_shape:
        try:
            tf = json.load(open(TRAFFIC_FILE))
            if bool(tf.get("free_surface_folded", False)) == args.fold:
                traffic = tf.get("hbm_bytes_per_launch")
        except Exception:
            traffic = None
    return {
        "value": npt * args.steps / wall / 1e9,
        "ms_per_step": wall / args.steps * 1e3,
        "kernel_ms": sum(kern_ms) / len(kern_ms),
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic},
        "npt": npt, "device": info, "dt": float(dt), "tuned": tuned,
        "countable": not args.fold and not args.contract and args.kernel == "auto", "shape_args": shape_args,
        "l1_last": {"step": it_last, "time_s": t_last, "l1_vs_analytical": l1_last},
        "pointwise_ref": None if pw_ms is None else {
            "kernel": "k_update_pressure16", "bytes_per_point": 20, "ms": pw_ms,
            "GBps": 20 * npt / (pw_ms / 1e3) / 1e9},
    }


def host_cpu():
    """CPU model, physical cores and logical CPUs of this host (lscpu, /proc/cpuinfo)."""
    import subprocess

    model, cores, logical = "unknown", None, os.cpu_count()
    try:
        info = {}
        for line in subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout.splitlines():
            k, _, v = line.partition(":")
            info[k.strip()] = v.strip()
        model = info.get("Model name", model)
        if info.get("Core(s) per socket") and info.get("Socket(s)"):
            cores = int(info["Core(s) per socket"]) * int(info["Socket(s)"])
    except Exception:
        pass
    if model == "unknown":
        try:
            for line in open("/proc/cpuinfo"):
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
        except OSError:
            pass
    return {"model": model, "physical_cores": cores, "logical_cpus": logical}


def run_cpu_harness(n, order, nt, ntry, fp64=False, threads=None, timeout=900):
    """oracle/_build/hpcscan_cpu[_fp64] -testCase Propa -testMode CacheBlk with
    the headline command line (eigen mode 2,2,2), OMP_PROC_BIND=close and
    OMP_PLACES=cores; returns (Best Gpoint/s eff., errTestCase, seconds, args)."""
    exe = os.path.join(ROOT, "oracle", "_build", "hpcscan_cpu_fp64" if fp64 else "hpcscan_cpu")
    if not os.path.exists(exe):
        raise RuntimeError(f"{exe} is missing: build the oracle (make -C oracle)")
    env = dict(os.environ, OMP_PROC_BIND="close", OMP_PLACES="cores")
    if threads:
        env["OMP_NUM_THREADS"] = str(threads)
    return run_harness(exe, "CacheBlk", n, order, nt, ntry, env, timeout)


def run_harness(exe, mode, n, order, nt, ntry, env=None, timeout=900):
    """exe -testCase Propa -testMode mode with the headline command line (eigen
    mode 2,2,2); returns (Best Gpoint/s eff., errTestCase, seconds, args)."""
    import subprocess

    cmd = [exe, "-testCase", "Propa", "-testMode", mode, "-propagator", "Ac2Standard", "-dim", "3",
           "-n1", str(n), "-n2", str(n), "-n3", str(n), "-fdOrder", str(order), "-nt", str(nt),
           "-ntry", str(ntry), "-param1", "2", "-param2", "2", "-param3", "2"]
    t0 = time.perf_counter()
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=timeout, env=env or dict(os.environ), cwd="/tmp")
    wall = time.perf_counter() - t0
    if r.returncode != 0:
        raise RuntimeError(f"{' '.join(cmd)} failed ({r.returncode}): {r.stdout[-2000:]}{r.stderr[-2000:]}")
    val = err = None
    for line in r.stdout.splitlines():
        t = line.strip()
        if t.startswith("Best Gpoint/s eff."):
            val = float(t.split()[-1])
        elif t.startswith("errTestCase2") or t.startswith("errTestCase"):
            err = float(t.split()[-1])
    if val is None:
        raise RuntimeError(f"no 'Best Gpoint/s eff.' line in the output of {' '.join(cmd)}")
    return val, err, wall, " ".join(cmd[1:])


def traffic_leg(args, shape_args):
    """HBM bytes per fused-step launch, measured in this run: two rocprofv3
    --pmc passes (FETCH_SIZE, then WRITE_SIZE; KiB per dispatch) over a short
    child run of this bench (same grid and shape, 1 + 3 steps, --calibrate-copy),
    each scaled by what the same counter reads for the plain copy of one whole
    grid (npoint x 4 bytes read and written) in that run; shape_args repeat the
    tile shape the timed launches used.  Returns None (the committed figure is
    used) when rocprofv3 gives no counter file."""
    import collections
    import csv
    import shutil
    import subprocess
    import tempfile

    from hpcscan_amd.layout import make_layout

    p = make_layout(3, (args.n, args.n, args.n), args.order)
    copy_kib = p.npoint * 4 / 1024.0
    child = [sys.executable, os.path.abspath(__file__), "--steps", "3", "--warmup", "1", "--n", str(args.n),
             "--order", str(args.order), "--kernel", args.kernel, "--no-cpu-baseline", "--no-adjoint", "--no-batch",
             "--no-autotune", "--no-analytical", "--no-pointwise-ref", "--no-traffic", "--calibrate-copy"] + shape_args
    got = {}
    for counter in ("FETCH_SIZE", "WRITE_SIZE"):
        d = tempfile.mkdtemp(prefix="hsx_pmc_", dir="/tmp")
        try:
            cmd = ["rocprofv3", "--pmc", counter, "--output-format", "csv", "-d", d, "-o", "p", "--"] + child
            t0 = time.perf_counter()
            try:
                r = subprocess.run(cmd, capture_output=True, text=True, timeout=150, cwd="/tmp",
                                   env=dict(os.environ, TMPDIR="/tmp"))
                rc = r.returncode
            except subprocess.TimeoutExpired:
                rc = "timeout"
            f = os.path.join(d, "p_counter_collection.csv")
            if not os.path.exists(f):
                log(f"traffic: rocprofv3 --pmc {counter} left no counter file ({rc}); using the committed figure")
                return None
            vals = collections.defaultdict(list)
            for row in csv.DictReader(open(f)):
                if row.get("Counter_Name") != counter:
                    continue
                k = row["Kernel_Name"]
                if "k_tiled<" in k:
                    vals["step"].append(float(row["Counter_Value"]))
                elif "k_array16" in k:
                    vals["copy"].append(float(row["Counter_Value"]))
            if not vals["step"] or len(vals["copy"]) != 1 or vals["copy"][0] <= 0:
                log(f"traffic: {counter} rows missing (step {len(vals['step'])}, copy {len(vals['copy'])})")
                return None
            got[counter] = (sum(vals["step"]) / len(vals["step"]), vals["copy"][0])
            log(f"traffic: {counter} {got[counter][0]:.1f} KiB per step launch, copy {got[counter][1]:.1f} KiB "
                f"({time.perf_counter() - t0:.0f} s pass)")
        finally:
            shutil.rmtree(d, ignore_errors=True)
    (fs, fc), (ws, wc) = got["FETCH_SIZE"], got["WRITE_SIZE"]
    fscale, wscale = copy_kib / fc, copy_kib / wc
    return {"bytes": (fs * fscale + ws * wscale) * 1024.0, "read_bytes": fs * fscale * 1024.0,
            "write_bytes": ws * wscale * 1024.0, "fetch_scale": round(fscale, 4), "write_scale": round(wscale, 4),
            "source": "rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE in this run, scaled by a whole-grid copy"}


def analytical_leg(args, cpu, gpu):
    """The reference's analytical-solution error on this workload: the MI355X
    harness (hpcscan_amd/bin/hpcscan -testMode MI355X) runs the CPU baseline's
    command line (checks at every snapInc-th step of the first try) and its
    errTestCase2 line is set beside the CPU path's; plus the timed run's own
    relative L1 error against the eigenmode at its last step."""
    exe = os.path.join(ROOT, "hpcscan_amd", "bin", "hpcscan")
    if not os.path.exists(exe):
        raise RuntimeError(f"{exe} is missing: build the harness (__graft_entry__.build())")
    _, err, wall, cmd = run_harness(exe, "MI355X", args.n, args.order, args.cpu_nt, 1, timeout=600)
    out = {"command": "hpcscan " + cmd, "errTestCase_mi355x": err,
           "errTestCase_cpu": None if cpu is None else cpu.get("errTestCase"),
           "timed_wavefield": gpu["l1_last"]}
    if out["errTestCase_cpu"] is not None and err is not None:
        out["agree_3_digits"] = f"{err:.3g}" == f"{out['errTestCase_cpu']:.3g}"
    log(f"analytical error: errTestCase2 MI355X {err} / CPU {out['errTestCase_cpu']} ({wall:.1f} s run)")
    return out


def cpu_leg(args):
    """The reference's CPU path restated (oracle/_build/hpcscan_cpu, testMode
    CacheBlk, cb 9999/4/16 as src/config.cpp:34-36) on this host's cores: the
    headline command line with a bounded number of steps (args.cpu_nt, one
    try).  value = its "Best Gpoint/s eff." report line."""
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or None
    hc = host_cpu()
    # two tries: the harness checks the error at every step of the first try
    # only and reports the best try, i.e. a check-free one
    val, err, wall, cmd = run_cpu_harness(args.n, args.order, args.cpu_nt, 2, threads=threads)
    cores = threads or os.cpu_count() or 1
    log(f"cpu baseline: {val:.3f} GPoint/s (best of 2 x {args.cpu_nt} steps, {wall:.1f} s run) on {cores} threads, "
        f"{hc['model']}")
    return {"value": val, "unit": "GPoint/s", "cores": cores, "kind": "port",
            "sample": f"hpcscan_cpu {cmd} (CacheBlk restatement, best of two tries of {args.cpu_nt} steps of the "
                      f"{args.n}^3 workload), {cores} OpenMP threads, OMP_PROC_BIND=close, OMP_PLACES=cores",
            "cpu_model": hc["model"], "physical_cores": hc["physical_cores"], "logical_cpus": hc["logical_cpus"],
            "errTestCase": err}


def adjoint_leg(args):
    """Misfit gradient of one shot: n^3 FP32 order 8, smooth velocity 2000-4000 m/s,
    h = 10 m, 8 Hz Ricker one point below the top centre, 200 receivers on a line
    one point below the top, observed data from a +-3 % smoothly perturbed model,
    nt steps at 0.8 x CFL.  Times (s per shot, whole gradient incl. forward):
      ours   hsx_gradient, checkpoints every K steps in the chosen 16-bit format
      plain  the same scheme in PyTorch tensor ops, FP32, autograd (full storage)
    Errors: relative L2 against the FP64 gradient (hsx_gradient on FP64 grids
    with storage-type checkpoints, i.e. exact FP64 recomputation).  Gradients
    run from a plan (work buffers allocated once per geometry, outside the
    timing); the *_allocating figures allocate inside the call."""
    import numpy as np
    import torch

    from hpcscan_amd import acquisition as acq
    from hpcscan_amd import capi
    from hpcscan_amd.layout import make_layout
    from hpcscan_amd.plain import plain_forward

    n, nt, K, order = args.adj_n, args.adj_nt, args.adj_every, args.adj_order
    fmt = {"fp16": capi.HSX_CKPT_FP16, "bf16": capi.HSX_CKPT_BF16, "fp32": capi.HSX_CKPT_FP32}[args.adj_format]
    h, vmin, vmax, f0 = 10.0, 2000.0, 4000.0, 8.0
    lay = {}
    for fp64 in (False, True):
        p = make_layout(3, (n, n, n), order, fp64=fp64)
        p.d = [h, h, h]
        lay[fp64] = p
    p = lay[False]
    v = acq.smooth_velocity(p, vmin, vmax, seed=4)
    m = 1.0 / (v * v)
    m_true = m * (1.0 + 0.03 * np.clip(acq.gaussian_field((p.n3, p.n2, p.n1), 44, 16.0).ravel() / 3.0, -1, 1))
    dt = float(np.float32(0.8 * acq.fd_cfl(order) * h / vmax))
    r = p.inner()
    inner = np.zeros((p.n3, p.n2, p.n1), bool)
    inner[r[4]:r[5] + 1, r[2]:r[3] + 1, r[0]:r[1] + 1] = True
    inner = inner.ravel()
    nshots = max(1, args.adj_shots)
    if nshots == 1:
        srcs = [acq.source_top_center(p, 1)]
    else:   # sources spread along n2 one point below the top, centre of n3
        N2 = r[3] - r[2] + 1
        srcs = [acq.flat_index(p, r[0] + 1, r[2] + (k + 1) * N2 // (nshots + 1), (r[4] + r[5]) // 2)
                for k in range(nshots)]
    src = np.array(srcs[:1], np.int64)
    rec = acq.receiver_line(p, 200, 1)
    wav = acq.ricker(nt, dt, f0)
    log(f"adjoint: {n}^3 O{order}, nt {nt}, dt {dt:.3e} s, {len(rec)} receivers, checkpoints every {K} "
        f"({args.adj_format}), {nshots} shot(s)")

    def dev(a):
        return capi.DeviceArray.from_numpy(np.ascontiguousarray(a))

    def setup(fp64):
        dtp = np.float64 if fp64 else np.float32
        g = capi.hsx_grid(lay[fp64])
        coef = dev(np.where(inner, dt * dt / m, 0.0).astype(dtp))
        coef_t = dev(np.where(inner, dt * dt / m_true, 0.0).astype(dtp))
        return dict(g=g, coef=coef, coef_t=coef_t, dtype=dtp)

    def with_source(S, s_idx):
        ds, dr, dw = dev(np.array([s_idx], np.int64)), dev(rec), dev(wav.astype(S["dtype"]).reshape(nt, 1))
        S["keep"] = (ds, dr, dw)
        S["shot"] = capi.make_shot(nt, dt, (ds, 1), dw, (dr, len(rec)))

    def observed(S64):
        # observed data: FP64 forward of the perturbed model (FP32 runs use it rounded)
        d64 = capi.DeviceArray(nt * len(rec) * 8)
        capi.forward(S64["g"], S64["coef_t"], S64["shot"], d64)
        dobs64 = d64.download(np.float64).reshape(nt, len(rec))
        return d64, dobs64, dev(dobs64.astype(np.float32))

    S64, S32 = setup(True), setup(False)
    with_source(S64, srcs[0])
    with_source(S32, srcs[0])
    d64, dobs64, dobs32 = observed(S64)

    def grad_of(S, dobs, every, f, reps=1, alloc_each=False):
        """best-of-reps seconds of one gradient; with a plan made beforehand
        (the work buffers allocated once per geometry) unless alloc_each, where
        each call allocates and frees them (hsx_gradient)"""
        plan = None if alloc_each else capi.AdjointPlan(S["g"], nt, len(rec), every, f, autotune=args.autotune)
        best, gout, J = 1e30, None, None
        for _ in range(reps):
            gr = capi.DeviceArray.from_numpy(np.zeros(p.npoint))
            capi.sync()
            t0 = time.perf_counter()
            if plan is None:
                J = capi.gradient(S["g"], S["coef"], S["shot"], dobs, gr, every, f)
            else:
                J = plan.gradient(S["g"], S["coef"], S["shot"], dobs, gr)
            best = min(best, time.perf_counter() - t0)
            gout = gr.download(np.float64)
            gr.free()
        if plan is not None:
            plan.close()
        return best, gout, J

    t64, g64, J64 = grad_of(S64, d64, args.adj_fp64_every or K, capi.HSX_CKPT_FP32)
    ours_t, g_ours, J32 = grad_of(S32, dobs32, K, fmt, reps=3 if nshots == 1 else 1)
    # further shots: gradients and misfits summed over the shots (FP64 on the host)
    for s_idx in srcs[1:]:
        with_source(S64, s_idx)
        with_source(S32, s_idx)
        d64, _, dobs32 = observed(S64)
        t, g, J = grad_of(S64, d64, args.adj_fp64_every or K, capi.HSX_CKPT_FP32)
        t64, g64, J64 = t64 + t, g64 + g, J64 + J
        t, g, J = grad_of(S32, dobs32, K, fmt)
        ours_t, g_ours, J32 = ours_t + t, g_ours + g, J32 + J
        log(f"adjoint: shot at {s_idx} done ({t:.3f} s)")
    if nshots > 1:
        log(f"adjoint: {nshots} shots, {ours_t:.3f} s in all ({ours_t / nshots:.3f} s per shot), FP64 {t64:.3f} s")
        ours_t /= nshots
        t64 /= nshots
    nrm = np.linalg.norm(g64)
    err_ours = float(np.linalg.norm(g_ours - g64) / nrm)
    full_t = err_full = full_alloc_t = ours_alloc_t = None
    if nshots == 1:
        ours_alloc_t, _, _ = grad_of(S32, dobs32, K, fmt, alloc_each=True)
    if nshots == 1 and (nt + 2) * p.npoint * 4 < 150e9:      # every step kept: only where it fits in HBM
        full_t, g_full, _ = grad_of(S32, dobs32, 0, capi.HSX_CKPT_FP32, reps=2)
        err_full = float(np.linalg.norm(g_full - g64) / nrm)
        full_alloc_t, _, _ = grad_of(S32, dobs32, 0, capi.HSX_CKPT_FP32, alloc_each=True)
    full_txt = "skipped" if full_t is None else f"{full_t:.3f} s (err {err_full:.2e}; {full_alloc_t:.3f} s allocating)"
    alloc_txt = "" if ours_alloc_t is None else f"; {ours_alloc_t:.3f} s allocating"
    log(f"adjoint: ours {ours_t:.3f} s (err {err_ours:.2e}{alloc_txt}), full storage FP32 {full_txt}, "
        f"FP64 {t64:.3f} s")

    # plain PyTorch, FP32, autograd over the whole time loop
    plain_t, err_plain = None, None
    if not args.no_plain and nshots == 1:
        mt = torch.tensor(m, dtype=torch.float32, device="cuda", requires_grad=True)
        dobs_t = torch.tensor(dobs64.astype(np.float32), device="cuda")
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        tr = plain_forward(p, mt, dt, src, wav.astype(np.float32), rec)
        loss = 0.5 * ((tr - dobs_t) ** 2).sum()
        loss.backward()
        torch.cuda.synchronize()
        plain_t = time.perf_counter() - t0
        gp = mt.grad.detach().double().cpu().numpy()
        err_plain = float(np.linalg.norm(gp - g64) / nrm)
        del mt, tr, loss
        torch.cuda.empty_cache()
        log(f"adjoint: plain PyTorch {plain_t:.3f} s (err {err_plain:.2e})")
    return {
        "workload": (f"one shot misfit gradient" if nshots == 1 else f"misfit gradient summed over {nshots} shots "
                     f"(sources along n2), one after the other") +
                    f", {n}^3 FP32 FD order {order}, {nt} steps, h 10 m, smooth velocity "
                    f"2000-4000 m/s, 8 Hz Ricker top centre, {len(rec)} receivers, data from a +-3% perturbed model",
        "s_per_shot": round(ours_t, 4),
        "s_per_shot_allocating": None if ours_alloc_t is None else round(ours_alloc_t, 4),
        "checkpoints": f"every {K} steps, {args.adj_format} with per-snapshot max-abs scale, FP64 gradient",
        "grad_rel_l2_vs_fp64": err_ours,
        "plain_s_per_shot": None if plain_t is None else round(plain_t, 4),
        "plain": "same scheme in PyTorch tensor ops, FP32, autograd tape (every step stored)",
        "plain_grad_rel_l2_vs_fp64": err_plain,
        "vs_plain": None if plain_t is None else round(plain_t / ours_t, 3),
        "full_storage_fp32_s_per_shot": None if full_t is None else round(full_t, 4),
        "full_storage_fp32_s_per_shot_allocating": None if full_alloc_t is None else round(full_alloc_t, 4),
        "vs_full_storage": None if full_t is None else round(full_t / ours_t, 3),
        "full_storage_fp32_grad_rel_l2_vs_fp64": err_full,
        "fp64_reference_s": round(t64, 4),
        "misfit_fp32": J32, "misfit_fp64": J64,
    }


def batch_leg(args):
    """Batched forward shots: args.batch_shots shots, cubes of N^3 inner points,
    N drawn (seeded) in [192, 320] (multiples of 4), FP32 order 8, h = 10 m,
    layered velocity 1500-4500 m/s with depth times (1 + 0.05 G), 10 Hz Ricker
    one point below the top centre, receivers along n2 one point below the
    top, nt steps at 0.8 x CFL.  All shots in one hsx_batch (one launch of each
    kernel per step, the time loop replayed from a HIP graph), against the same
    shots one after the other through hsx_forward."""
    import numpy as np

    from hpcscan_amd import acquisition as acq
    from hpcscan_amd import capi
    from hpcscan_amd.layout import make_layout

    rng = np.random.default_rng(2024)
    ns, nt, h, vmin, vmax, f0 = args.batch_shots, args.batch_nt, 10.0, 1500.0, 4500.0, 10.0
    sizes = [int(4 * round(x / 4)) for x in rng.uniform(192, 320, ns)]
    dt = float(np.float32(0.8 * acq.fd_cfl(8) * h / vmax))
    wav = acq.ricker(nt, dt, f0).astype(np.float32).reshape(nt, 1)
    grids, keep, shots, traces, pts = [], [], [], [], 0
    t_setup = time.perf_counter()
    for s, n in enumerate(sizes):
        p = make_layout(3, (n, n, n), 8)
        p.d = [h, h, h]
        v = acq.layered_velocity(p, vmin, vmax, 0.05, seed=100 + s)
        r = p.inner()
        inner = np.zeros((p.n3, p.n2, p.n1), bool)
        inner[r[4]:r[5] + 1, r[2]:r[3] + 1, r[0]:r[1] + 1] = True
        coef = capi.DeviceArray.from_numpy(np.where(inner.ravel(), (v * dt) ** 2, 0.0).astype(np.float32))
        src = capi.DeviceArray.from_numpy(np.array([acq.source_top_center(p, 1)], np.int64))
        rec_h = acq.receiver_line(p, None, 1)
        rec = capi.DeviceArray.from_numpy(rec_h)
        w = capi.DeviceArray.from_numpy(wav)
        tr = capi.DeviceArray(nt * len(rec_h) * 4)
        shot = capi.make_shot(nt, dt, (src, 1), w, (rec, len(rec_h)))
        grids.append(capi.hsx_grid(p))
        keep.append((coef, src, rec, w, tr, len(rec_h)))
        shots.append(shot)
        traces.append(tr)
        pts += n ** 3
    log(f"batch: {ns} shots, N in [{min(sizes)}, {max(sizes)}], {pts / 1e9:.2f} G inner points, "
        f"setup {time.perf_counter() - t_setup:.1f} s")
    b = capi.Batch(grids, [k[0] for k in keep], shots, traces)
    b.run(use_graph=True)                   # capture + first replay
    t0 = time.perf_counter()
    b.run(use_graph=True)
    t_default = time.perf_counter() - t0
    shape, sms = b.autotune(reps=3)
    log(f"batch: default shape {t_default:.3f} s; autotuned shape {shape} ({sms:.3f} ms per step in the sweep)")
    b.run(use_graph=True)                   # capture the tuned step + first replay
    t0 = time.perf_counter()
    b.run(use_graph=True)
    t_graph = time.perf_counter() - t0
    t0 = time.perf_counter()
    b.run(use_graph=False)
    t_nograph = time.perf_counter() - t0
    batch_tr = [k[4].download(np.float32) for k in keep[:2]]
    b.close()
    # the same shots one by one
    t0 = time.perf_counter()
    for g, k, shot in zip(grids, keep, shots):
        capi.forward(g, k[0], shot, k[4])
    t_single = time.perf_counter() - t0
    same = all(np.array_equal(batch_tr[i], keep[i][4].download(np.float32)) for i in range(len(batch_tr)))
    log(f"batch: graph {t_graph:.3f} s, no graph {t_nograph:.3f} s, one by one {t_single:.3f} s, same={same}")
    return {
        "workload": f"{ns} forward shots, N^3 with N in [{min(sizes)}, {max(sizes)}], FP32 order 8, {nt} steps, "
                    f"layered 1500-4500 m/s x (1 + 0.05 G), 10 Hz Ricker",
        "gpoint_per_s": round(pts * nt / t_graph / 1e9, 3),
        "roofline_frac": round(16 * pts * nt / t_graph / 1e9 / HBM_PEAK_GBS, 4),
        "s_graph": round(t_graph, 4), "s_no_graph": round(t_nograph, 4), "s_one_by_one": round(t_single, 4),
        "s_graph_default_shape": round(t_default, 4),
        "tile": {"lanes": shape[0], "rows": shape[1], "chunk": shape[2], "ring": shape[3]},
        "traces_equal_single_shot": bool(same),
    }


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--n", type=int, default=1000, help="inner points per axis")
    ap.add_argument("--order", type=int, default=8, choices=[4, 8, 12, 16])
    ap.add_argument("--kernel", default="auto", choices=["auto", "ref", "tiled"])
    ap.add_argument("--cpu-nt", type=int, default=25,
                    help="time steps of each of the two tries of the bounded CPU-baseline sample "
                         "(about 20 s at 1000^3 on 16 cores)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-analytical", action="store_true",
                    help="skip the MI355X harness run that reports errTestCase2 on the CPU baseline's command line")
    ap.add_argument("--tile-lx", type=int, default=0, choices=[0, 16, 32, 64, 128],
                    help="lanes along n1 per tile of the fused kernel (0 = chosen from the grid width)")
    ap.add_argument("--tile-rows", type=int, default=0, choices=[0, 8, 16],
                    help="rows along n2 per tile (0 = library choice)")
    ap.add_argument("--tile-chunk", type=int, default=0,
                    help="n3 planes streamed per workgroup (0 = library choice)")
    ap.add_argument("--streams", default="default", choices=["default", "cached", "nt"],
                    help="cached or nontemporal prn/coef/output accesses of the fused kernel")
    ap.add_argument("--fold", action="store_true",
                    help="free surface folded into the fused step (hsx_ac2_step_fs) instead of its own launches")
    ap.add_argument("--no-ring", action="store_true",
                    help="load each plane's halo one plane ahead instead of with the queue front")
    ap.add_argument("--contract", action="store_true",
                    help="fused multiply-adds in the fused step (hsx_set_contract): not the reference's bits")
    ap.add_argument("--no-autotune", dest="autotune", action="store_false",
                    help="keep the library's default tile shape (no hsx_autotune before timing)")
    ap.add_argument("--no-traffic", dest="traffic", action="store_false",
                    help="take roofline.traffic from the committed counter summary instead of measuring it with "
                         "two rocprofv3 --pmc child runs")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="run one full-grid copy kernel first (counter calibration for profiles)")
    ap.add_argument("--no-batch", dest="batch", action="store_false",
                    help="skip the batched-shots leg (32 uneven shots)")
    ap.add_argument("--batch-shots", type=int, default=32)
    ap.add_argument("--batch-nt", type=int, default=1000)
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint-gradient leg")
    ap.add_argument("--no-pointwise-ref", dest="pointwise_ref", action="store_false",
                    help="skip the pointwise streaming reference (updatePressure on the same grid)")
    ap.add_argument("--no-plain", action="store_true", help="adjoint leg without the plain PyTorch timing")
    ap.add_argument("--adj-n", type=int, default=200)
    ap.add_argument("--adj-nt", type=int, default=1000)
    ap.add_argument("--adj-order", type=int, default=8, choices=[4, 8, 12, 16])
    ap.add_argument("--adj-fp64-every", type=int, default=0,
                    help="checkpoint interval of the FP64 reference gradient (0 = the same as --adj-every)")
    ap.add_argument("--adj-every", type=int, default=20)
    ap.add_argument("--adj-shots", type=int, default=1,
                    help="shots whose gradients are summed (one after the other: 8 shots of 512^3 x 2000 steps "
                         "need ~0.6 TB of checkpoints together)")
    ap.add_argument("--adj-format", default="fp16", choices=["fp16", "bf16", "fp32"])
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(f"bench.py measures one MI355X: one process on one GPU (--gpus {args.gpus}, WORLD_SIZE "
              f"{world} requested). Sharding the stencil over GPUs (slabs along n3 with one halo "
              f"exchange per step) is not part of this library.",
              file=sys.stderr)
        sys.exit(2)

    gpu = gpu_leg(args)
    adj = None if args.no_adjoint else adjoint_leg(args)
    bat = batch_leg(args) if args.batch else None
    cpu = None if args.no_cpu_baseline else cpu_leg(args)
    # the child run repeats this run's tile shape; the folded free surface and the
    # contracted arithmetic are not counted
    if args.traffic and gpu["countable"]:
        tr = None
        try:
            tr = traffic_leg(args, gpu["shape_args"])
        except Exception as e:
            log(f"traffic: measurement failed ({e}); using the committed figure")
        if tr is not None:
            gpu["roofline"]["traffic"] = tr["bytes"]
            gpu["roofline"]["traffic_measured"] = tr
    ana = None
    if not args.no_analytical:
        try:
            ana = analytical_leg(args, cpu, gpu)
        except Exception as e:       # the headline line is printed anyway, with the failure
            log(f"analytical error: the MI355X harness run failed: {e}")
            ana = {"error": str(e)[-500:], "timed_wavefield": gpu["l1_last"]}
    npt = gpu["npt"]
    pts_stencil = 6 * (args.order // 2) + 1
    out = {
        "metric": METRIC,
        "value": round(gpu["value"], 3),
        "unit": "GPoint/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(gpu["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"Propa Ac2Standard FP32 FD order {args.order}, {args.n}^3 inner grid, "
                               f"eigen mode (2,2,2), free surface on all faces, dt at the CFL limit",
                   "grid": [args.n] * 3, "fd_order": args.order, "propagator": "Ac2Standard",
                   "inner_points": npt, "kernel": args.kernel,
                   "tile_lx": args.tile_lx, "tile_rows": args.tile_rows, "tile_chunk": args.tile_chunk,
                   "streams": args.streams, "halo_ring": not args.no_ring,
                   "free_surface": "folded into the step" if args.fold else "own launches before the step",
                   "arithmetic": "contracted (fma)" if args.contract else "reference order, no fma",
                   "autotuned_tile": gpu["tuned"]},
        "roofline": {k: (round(v, 4) if isinstance(v, float) else v) for k, v in gpu["roofline"].items()},
        "cpu_baseline": cpu,
        "analytical_error": ana,
        "adjoint": adj,
        "batch": bat,
        "extra": {"fused_kernel_ms": round(gpu["kernel_ms"], 4),
                  "hpcscan_apparent_GBps": round(gpu["value"] * (pts_stencil + 4) * 4, 1),
                  "device": gpu["device"]["name"], "arch": gpu["device"]["arch"],
                  "pointwise_ref": gpu["pointwise_ref"]},
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
