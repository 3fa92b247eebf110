"""Shot modeling and the misfit gradient on the GPU against the CPU oracle:
traces and gradients bit for bit (full storage and storage-type checkpoints),
16-bit checkpoints against the FP64 oracle gradient, and the building blocks
(sources / receivers, imaging condition, checkpoint codec, out-of-place step)."""
import numpy as np
import pytest

from hpcscan_amd.layout import make_layout
from tests.shots import make_shot

pytestmark = pytest.mark.gpu


def _same(a, b):
    assert a.dtype == b.dtype and a.shape == b.shape
    eq = (a == b) | (np.isnan(a) & np.isnan(b))
    if not eq.all():
        i = np.flatnonzero(~eq.ravel())[:5]
        raise AssertionError(f"{(~eq).sum()} values differ, e.g. at {i}: {a.ravel()[i]} vs {b.ravel()[i]}")


def _dev(capi, a):
    return capi.DeviceArray.from_numpy(np.ascontiguousarray(a))


def _shot_dev(capi, sh):
    d = dict(src=_dev(capi, sh["src"]), rec=_dev(capi, sh["rec"]), wav=_dev(capi, sh["wavelet"]),
             coef=_dev(capi, sh["coef"]))
    s = capi.make_shot(sh["nt"], sh["dt"], (d["src"], len(sh["src"])), d["wav"], (d["rec"], len(sh["rec"])))
    return d, s


SIZES = [(20, 18, 16), (68, 40, 24)]   # narrow tiles (8 rows) / 32-lane 16-row tiles with the halo ring


@pytest.mark.parametrize("n", SIZES)
@pytest.mark.parametrize("fp64", [False, True])
@pytest.mark.parametrize("order", [4, 8])
def test_forward_traces_bit_exact(oracle, capi, n, fp64, order):
    print(f"forward n={n} fp64={fp64} order={order}", flush=True)
    sh = make_shot(oracle, n, order, nt=40, fp64=fp64, seed=order, with_obs=False)
    g = capi.hsx_grid(sh["p"])
    d, s = _shot_dev(capi, sh)
    tr = capi.DeviceArray(sh["nt"] * len(sh["rec"]) * np.dtype(sh["dtype"]).itemsize)
    capi.forward(g, d["coef"], s, tr)
    ref = oracle.forward(sh["o"], sh["coef"], sh["iv"], sh["nt"], sh["src"], sh["wavelet"], sh["rec"])
    assert np.abs(ref).max() > 0
    _same(tr.download(sh["dtype"]).reshape(ref.shape), ref)


@pytest.mark.parametrize("fp64,every", [(False, 0), (False, 7), (True, 0), (True, 9), (False, 45)])
def test_gradient_bit_exact(oracle, capi, fp64, every):
    """full storage and storage-type checkpoints give the oracle's gradient bit for bit"""
    print(f"gradient fp64={fp64} every={every}", flush=True)
    sh = make_shot(oracle, (24, 20, 18), 8, nt=45, fp64=fp64, seed=11)
    g = capi.hsx_grid(sh["p"])
    d, s = _shot_dev(capi, sh)
    dobs = _dev(capi, sh["d_obs"])
    grad = _dev(capi, np.zeros(sh["p"].npoint))
    J = capi.gradient(g, d["coef"], s, dobs, grad, every, capi.HSX_CKPT_FP32)
    gref, Jref = oracle.gradient(sh["o"], sh["coef"], sh["iv"], sh["dt"], sh["nt"], sh["src"], sh["wavelet"],
                                 sh["rec"], sh["d_obs"])
    assert J == pytest.approx(Jref, rel=1e-12)
    assert np.abs(gref).max() > 0
    _same(grad.download(np.float64), gref)


@pytest.mark.parametrize("nsrc", [2, 70])
@pytest.mark.parametrize("fp64", [False, True])
def test_receivers_on_source_points(oracle, capi, nsrc, fp64):
    """a whole plane of receivers (1152, several gather blocks) through which
    the sources pass: the receivers on source points record the field after
    the injection, as in the oracle; 2 sources take the one-launch source and
    receiver kernel, 70 the separate inject and gather launches.  Forward
    traces and the full-storage gradient (reverse source term and residual
    in one launch) equal the oracle's bit for bit."""
    print(f"receivers on sources nsrc={nsrc} fp64={fp64}", flush=True)
    sh = make_shot(oracle, (40, 36, 32), 8, nt=24, fp64=fp64, seed=31)
    p, r = sh["p"], sh["p"].inner()
    i1 = r[0] + 2
    rec = np.array([i1 + i2 * p.n1 + i3 * p.n1 * p.n2 for i3 in range(r[4], r[5] + 1)
                    for i2 in range(r[2], r[3] + 1)], np.int64)
    rng = np.random.default_rng(5)
    src = np.concatenate([sh["src"], rng.choice(rec, nsrc - 1, replace=False)])
    assert np.isin(src, rec).all() and len(np.unique(src)) == nsrc
    wav = (sh["wavelet"] * rng.uniform(0.5, 1.5, nsrc)).astype(sh["dtype"])
    o, coef, iv, nt = sh["o"], sh["coef"], sh["iv"], sh["nt"]
    coef_true = sh["coef_true"]
    d_obs = oracle.forward(o, coef_true, iv, nt, src, wav, rec)
    ref = oracle.forward(o, coef, iv, nt, src, wav, rec)
    g = capi.hsx_grid(p)
    dsrc, drec, dwav, dcoef = (_dev(capi, x) for x in (src, rec, wav, coef))
    s = capi.make_shot(nt, sh["dt"], (dsrc, nsrc), dwav, (drec, len(rec)))
    tr = capi.DeviceArray(nt * len(rec) * np.dtype(sh["dtype"]).itemsize)
    capi.forward(g, dcoef, s, tr)
    _same(tr.download(sh["dtype"]).reshape(ref.shape), ref)
    grad = _dev(capi, np.zeros(p.npoint))
    J = capi.gradient(g, dcoef, s, _dev(capi, d_obs), grad, 0, capi.HSX_CKPT_FP32)
    gref, Jref = oracle.gradient(o, coef, iv, sh["dt"], nt, src, wav, rec, d_obs)
    assert J == pytest.approx(Jref, rel=1e-12)
    assert np.abs(gref).max() > 0
    _same(grad.download(np.float64), gref)


@pytest.mark.parametrize("fmt,tol", [("fp16", 1e-3), ("bf16", 1e-3)])
def test_low_precision_checkpoints_against_fp64_oracle(oracle, capi, fmt, tol):
    """FP32 stencil, 16-bit scaled checkpoints, FP64 gradient accumulation: the
    error against the FP64 oracle gradient within 1e-3 relative L2 for both
    16-bit formats"""
    sh64 = make_shot(oracle, (40, 36, 32), 8, nt=120, fp64=True, seed=21)
    sh = make_shot(oracle, (40, 36, 32), 8, nt=120, fp64=False, seed=21)
    g64, _ = oracle.gradient(sh64["o"], sh64["coef"], sh64["iv"], sh64["dt"], sh64["nt"], sh64["src"],
                             sh64["wavelet"], sh64["rec"], sh64["d_obs"])
    g = capi.hsx_grid(sh["p"])
    d, s = _shot_dev(capi, sh)
    dobs = _dev(capi, sh["d_obs"])
    errs = {}
    for name, f in (("fp32", capi.HSX_CKPT_FP32), (fmt, {"fp16": capi.HSX_CKPT_FP16, "bf16": capi.HSX_CKPT_BF16}[fmt])):
        grad = _dev(capi, np.zeros(sh["p"].npoint))
        capi.gradient(g, d["coef"], s, dobs, grad, 20, f)
        x = grad.download(np.float64)
        errs[name] = float(np.linalg.norm(x - g64) / np.linalg.norm(g64))
    print(f"relative L2 error vs FP64 oracle: {errs}", flush=True)
    assert errs["fp32"] < 1e-4
    assert errs[fmt] < tol


@pytest.mark.parametrize("fmt", ["fp16", "bf16"])
def test_checkpoint_codec(capi, fmt):
    f = {"fp16": capi.HSX_CKPT_FP16, "bf16": capi.HSX_CKPT_BF16}[fmt]
    ulp = {"fp16": 2.0 ** -11, "bf16": 2.0 ** -8}[fmt]
    rng = np.random.default_rng(4)
    n = 100003
    u = (rng.standard_normal(n) * 3e-3).astype(np.float32)
    up = (u - rng.standard_normal(n).astype(np.float32) * 1e-6).astype(np.float32)
    du, dup = _dev(capi, u), _dev(capi, up)
    packed = capi.DeviceArray(capi.ckpt_bytes(n))
    out = capi.DeviceArray(n * 4)
    capi.ckpt_encode(du, None, n, packed, f)
    capi.ckpt_decode(packed, n, f, None, out)
    x = out.download(np.float32)
    big = np.abs(u) > np.abs(u).max() * 2.0 ** -13
    assert np.all(np.abs(x - u)[big] <= ulp * np.abs(u)[big] * 1.0001)
    # pair: u - u_prev keeps its own scale; decoding with base u gives u_prev back
    capi.ckpt_encode(du, dup, n, packed, f)
    capi.ckpt_decode(packed, n, f, du, out)
    y = out.download(np.float32)
    dd = (u - up).astype(np.float64)
    assert np.max(np.abs(y - up)) <= ulp * np.abs(dd).max() * 1.01 + 2 * np.spacing(np.abs(u).max())
    z = np.zeros(n, np.float32)
    capi.ckpt_encode(_dev(capi, z), None, n, packed, f)
    capi.ckpt_decode(packed, n, f, None, out)
    assert np.all(out.download(np.float32) == 0)


@pytest.mark.parametrize("mode", ["add", "set"])
def test_inject_and_gather(capi, mode):
    p = make_layout(3, (12, 11, 10), 4)
    g = capi.hsx_grid(p)
    rng = np.random.default_rng(2)
    u = rng.standard_normal(p.npoint).astype(np.float32)
    idx = rng.choice(p.npoint, 37, replace=False).astype(np.int64)
    val = rng.standard_normal(37).astype(np.float32)
    w = rng.standard_normal(37).astype(np.float32)
    du, di, dv, dw = (_dev(capi, x) for x in (u, idx, val, w))
    m = capi.HSX_INJECT_ADD if mode == "add" else capi.HSX_INJECT_SET
    capi.inject(g, du, 37, di, dv, dw, 0.5, m)
    ref = u.copy()
    v = 0.5 * val.astype(np.float64) * w.astype(np.float64)
    ref[idx] = (ref[idx].astype(np.float64) + v).astype(np.float32) if mode == "add" else v.astype(np.float32)
    got = du.download(np.float32)
    _same(got, ref)
    out = capi.DeviceArray(37 * 4)
    capi.gather(g, du, 37, di, out)
    _same(out.download(np.float32), got[idx])


def test_imaging_condition(capi):
    p = make_layout(3, (13, 12, 11), 8)
    g = capi.hsx_grid(p)
    rng = np.random.default_rng(3)
    mu, un, u, up = (rng.standard_normal(p.npoint).astype(np.float32) for _ in range(4))
    grad0 = rng.standard_normal(p.npoint)
    r = p.inner()
    rim = [r[0] + 1, r[1] - 1, r[2] + 1, r[3] - 1, r[4] + 1, r[5] - 1]
    src = rim[0] + 2 + (rim[2] + 3) * p.n1 + (rim[4] + 1) * p.n1 * p.n2
    dg = _dev(capi, grad0)
    capi.imaging(g, dg, *(_dev(capi, x) for x in (mu, un, u, up)), rim, -3.5, src, 0.25)
    ref = grad0.copy()
    box = np.zeros((p.n3, p.n2, p.n1), bool)
    box[rim[4]:rim[5] + 1, rim[2]:rim[3] + 1, rim[0]:rim[1] + 1] = True
    box = box.ravel()
    d2 = (un.astype(np.float64) - 2.0 * u.astype(np.float64)) + up.astype(np.float64)
    d2[src] -= 0.25
    ref[box] += -3.5 * mu.astype(np.float64)[box] * d2[box]
    _same(dg.download(np.float64), ref)


def test_out_of_place_step(oracle, capi):
    p = make_layout(3, (36, 20, 18), 8)
    g = capi.hsx_grid(p)
    rng = np.random.default_rng(9)
    prn, prc, coef = (rng.standard_normal(p.npoint).astype(np.float32) for _ in range(3))
    dprn, dprc, dcoef = (_dev(capi, x) for x in (prn, prc, coef))
    out = capi.DeviceArray(p.npoint * 4)
    capi.ac2_step_to(g, out, dprn, dprc, dcoef)
    capi.ac2_step(g, dprn, dprc, dcoef)
    _same(out.download(np.float32)[np.array(masks_inner(p))], dprn.download(np.float32)[np.array(masks_inner(p))])
    with pytest.raises(capi.HsxError):
        capi.ac2_step_to(g, dprc, dprn, dprc, dcoef)


def masks_inner(p):
    m = np.zeros((p.n3, p.n2, p.n1), bool)
    r = p.inner()
    m[r[4]:r[5] + 1, r[2]:r[3] + 1, r[0]:r[1] + 1] = True
    return m.ravel()


@pytest.mark.parametrize("fp64", [False, True])
@pytest.mark.parametrize("use_graph", [False, True])
@pytest.mark.parametrize("order", [8, 4])
def test_batched_shots_equal_single_shots(oracle, capi, fp64, use_graph, order):
    """uneven grids in one batch (one fused-step launch per time step, optionally
    a captured graph replayed twice): every shot's traces equal the oracle's.
    Order 8 (and FP64 order 4) takes the 16-byte n1-face kernel, FP32 order 4
    (R = 2, not a multiple of the vector) the per-point face kernels"""
    sizes = [(36, 20, 18), (40, 24, 22), (28, 30, 26), (68, 36, 20)]
    shots, keep = [], []
    for k, n in enumerate(sizes):
        sh = make_shot(oracle, n, order, nt=30, fp64=fp64, seed=50 + k, with_obs=False)
        d, s = _shot_dev(capi, sh)
        tr = capi.DeviceArray(sh["nt"] * len(sh["rec"]) * np.dtype(sh["dtype"]).itemsize)
        shots.append((sh, d, s, tr))
    b = capi.Batch([capi.hsx_grid(x[0]["p"]) for x in shots], [x[1]["coef"] for x in shots],
                   [x[2] for x in shots], [x[3] for x in shots])
    for rep in range(2 if use_graph else 1):
        b.run(use_graph=use_graph)
        for sh, d, s, tr in shots:
            ref = oracle.forward(sh["o"], sh["coef"], sh["iv"], sh["nt"], sh["src"], sh["wavelet"], sh["rec"])
            _same(tr.download(sh["dtype"]).reshape(ref.shape), ref)
    b.close()


@pytest.mark.parametrize("fp64", [False, True])
def test_batched_shots_several_rounds(oracle, capi, fp64):
    """a batch whose tiles outnumber the resident workgroups (16 shots of 800
    rows: 50 tile rows each), so the launch streams short n3 chunks in several
    rounds, with uneven last chunks (N3 = 36..66): traces equal the oracle's"""
    sizes = [(16 + 4 * (k % 3), 800, 36 + 2 * k) for k in range(16)]
    shots = []
    for k, n in enumerate(sizes):
        sh = make_shot(oracle, n, 8, nt=10, fp64=fp64, seed=90 + k, with_obs=False)
        d, s = _shot_dev(capi, sh)
        tr = capi.DeviceArray(sh["nt"] * len(sh["rec"]) * np.dtype(sh["dtype"]).itemsize)
        shots.append((sh, d, s, tr))
    print(f"batch of {len(sizes)} shots, fp64={fp64}", flush=True)
    b = capi.Batch([capi.hsx_grid(x[0]["p"]) for x in shots], [x[1]["coef"] for x in shots],
                   [x[2] for x in shots], [x[3] for x in shots])
    b.run(use_graph=False)
    for sh, d, s, tr in shots:
        ref = oracle.forward(sh["o"], sh["coef"], sh["iv"], sh["nt"], sh["src"], sh["wavelet"], sh["rec"])
        _same(tr.download(sh["dtype"]).reshape(ref.shape), ref)
    b.close()


@pytest.mark.parametrize("fmt", ["fp16", "bf16"])
def test_16bit_checkpoints_odd_point_count(oracle, capi, fmt):
    """an odd number of grid points (41 x 37 x 33 inner, order 8: 49 x 45 x 41
    points): every snapshot after the first starts at an odd multiple of
    2 bytes unless the snapshot size is padded; the gradient is the one of
    aligned snapshots (against the FP64 oracle, and equal to storage-type
    checkpoints within the 16-bit error)"""
    n = (41, 37, 33)
    sh64 = make_shot(oracle, n, 8, nt=60, fp64=True, seed=23)
    sh = make_shot(oracle, n, 8, nt=60, fp64=False, seed=23)
    assert sh["p"].npoint % 2 == 1
    assert capi.ckpt_bytes(sh["p"].npoint) % 256 == 0
    g64, _ = oracle.gradient(sh64["o"], sh64["coef"], sh64["iv"], sh64["dt"], sh64["nt"], sh64["src"],
                             sh64["wavelet"], sh64["rec"], sh64["d_obs"])
    g = capi.hsx_grid(sh["p"])
    d, s = _shot_dev(capi, sh)
    dobs = _dev(capi, sh["d_obs"])
    f = {"fp16": capi.HSX_CKPT_FP16, "bf16": capi.HSX_CKPT_BF16}[fmt]
    grad = _dev(capi, np.zeros(sh["p"].npoint))
    capi.gradient(g, d["coef"], s, dobs, grad, 7, f)
    x = grad.download(np.float64)
    err = float(np.linalg.norm(x - g64) / np.linalg.norm(g64))
    print(f"{fmt} checkpoints every 7, odd npoint {sh['p'].npoint}: rel L2 vs FP64 oracle {err:.3e}", flush=True)
    assert np.all(x[~sh["grad_mask"]] == 0)
    assert err < 1e-3


def test_plan_and_given_residual(oracle, capi):
    """one plan serves two shots of one geometry; a residual passed in
    (HSX_ADJ_RESIDUAL) gives the gradient of the formed residual bit for bit"""
    sh = make_shot(oracle, (24, 20, 18), 8, nt=45, seed=27)
    g = capi.hsx_grid(sh["p"])
    d, s = _shot_dev(capi, sh)
    plan = capi.AdjointPlan(g, sh["nt"], len(sh["rec"]), 9, capi.HSX_CKPT_FP32)
    rplan = capi.AdjointPlan(g, sh["nt"], len(sh["rec"]), 9, capi.HSX_CKPT_FP32, residual=True)
    trace = oracle.forward(sh["o"], sh["coef"], sh["iv"], sh["nt"], sh["src"], sh["wavelet"], sh["rec"])
    for k, dobs_h in enumerate((sh["d_obs"], (sh["d_obs"] * 0.5).astype(np.float32))):
        gref, Jref = oracle.gradient(sh["o"], sh["coef"], sh["iv"], sh["dt"], sh["nt"], sh["src"], sh["wavelet"],
                                     sh["rec"], dobs_h)
        grad = _dev(capi, np.zeros(sh["p"].npoint))
        J = plan.gradient(g, d["coef"], s, _dev(capi, dobs_h), grad)
        assert J == pytest.approx(Jref, rel=1e-12)
        _same(grad.download(np.float64), gref)
        res = (trace - dobs_h).astype(np.float32)
        grad_r = _dev(capi, np.zeros(sh["p"].npoint))
        J_r = rplan.gradient(g, d["coef"], s, _dev(capi, res), grad_r)
        assert J_r == pytest.approx(Jref, rel=1e-12), k
        _same(grad_r.download(np.float64), gref)
    # a plan refuses another geometry
    sh2 = make_shot(oracle, (24, 20, 20), 8, nt=45, seed=27, with_obs=False)
    d2, s2 = _shot_dev(capi, sh2)
    with pytest.raises(capi.HsxError):
        plan.gradient(capi.hsx_grid(sh2["p"]), d2["coef"], s2, _dev(capi, sh["d_obs"]),
                      _dev(capi, np.zeros(sh2["p"].npoint)))
    plan.close()
    rplan.close()


def test_source_on_a_boundary_plane_bit_exact(oracle, capi):
    """sources on and off the n1 boundary plane: the term of a source the free
    surface zeroes is left out of the imaging condition, as in the oracle"""
    sh = make_shot(oracle, (24, 20, 18), 8, nt=40, seed=29)
    p = sh["p"]
    r = p.inner()
    mid2, mid3 = (r[2] + r[3]) // 2, (r[4] + r[5]) // 2
    sh["src"] = np.array([r[0] + mid2 * p.n1 + mid3 * p.n1 * p.n2, (r[0] + 3) + (mid2 + 1) * p.n1 + mid3 * p.n1 * p.n2],
                         np.int64)
    sh["wavelet"] = np.repeat(sh["wavelet"], 2, axis=1).copy()
    d_obs = oracle.forward(sh["o"], sh["coef_true"], sh["iv"], sh["nt"], sh["src"], sh["wavelet"], sh["rec"])
    g = capi.hsx_grid(p)
    d, s = _shot_dev(capi, sh)
    grad = _dev(capi, np.zeros(p.npoint))
    capi.gradient(g, d["coef"], s, _dev(capi, d_obs), grad, 0, capi.HSX_CKPT_FP32)
    gref, _ = oracle.gradient(sh["o"], sh["coef"], sh["iv"], sh["dt"], sh["nt"], sh["src"], sh["wavelet"], sh["rec"],
                              d_obs)
    x = grad.download(np.float64)
    assert np.all(x[~sh["grad_mask"]] == 0)
    _same(x, gref)


def test_indices_outside_the_grid_are_refused_on_the_host(capi):
    p = make_layout(3, (12, 11, 10), 4)
    capi.check_indices([0, p.npoint - 1], p.npoint)
    for bad in ([-1], [p.npoint], [3, p.npoint + 5]):
        with pytest.raises(ValueError):
            capi.check_indices(bad, p.npoint)


@pytest.mark.parametrize("fp64", [False, True])
def test_batch_autotune_keeps_results(oracle, capi, fp64):
    """hsx_batch_autotune times the candidate shapes of the batched step and
    keeps the fastest; the batch then still gives the oracle's traces, with
    and without the captured graph (captured again after the new shape)"""
    sizes = [(36, 20, 18), (68, 36, 20), (40, 24, 22)]
    shots = []
    for k, n in enumerate(sizes):
        sh = make_shot(oracle, n, 8, nt=24, fp64=fp64, seed=70 + k, with_obs=False)
        d, s = _shot_dev(capi, sh)
        tr = capi.DeviceArray(sh["nt"] * len(sh["rec"]) * np.dtype(sh["dtype"]).itemsize)
        shots.append((sh, d, s, tr))
    b = capi.Batch([capi.hsx_grid(x[0]["p"]) for x in shots], [x[1]["coef"] for x in shots],
                   [x[2] for x in shots], [x[3] for x in shots])
    b.run(use_graph=True)                   # capture with the first shape
    shape, ms = b.autotune(reps=2)
    print(f"batch autotune fp64={fp64}: shape {shape}, {ms:.4f} ms per step", flush=True)
    assert shape[0] in (32, 64) and shape[1] in (8, 16) and ms > 0
    for use_graph in (True, False):
        b.run(use_graph=use_graph)
        for sh, d, s, tr in shots:
            ref = oracle.forward(sh["o"], sh["coef"], sh["iv"], sh["nt"], sh["src"], sh["wavelet"], sh["rec"])
            _same(tr.download(sh["dtype"]).reshape(ref.shape), ref)
    b.close()


@pytest.mark.parametrize("fp64", [False, True])
def test_batch_persistent_step_bit_exact(oracle, capi, fp64):
    """the batched step as one persistent launch (chunk -1: each resident
    workgroup streams an equal share of the shots' whole tile columns, cut
    across column ends) with two tile shapes each (FP32: 64 x 16 ring, 32 x 8;
    FP64: 32 x 8, 64 x 8), set by
    hsx_batch_set_shape, and a batch with far more columns than workgroups
    (16 shots of 800 rows): traces equal the oracle's, with and without the
    graph; a shape that does not exist is refused"""
    for sizes, nt in (([(36, 20, 18), (68, 36, 20), (40, 24, 22)], 24),
                      ([(16 + 4 * (k % 3), 800, 36 + 2 * k) for k in range(16)], 10)):
        shots = []
        for k, n in enumerate(sizes):
            sh = make_shot(oracle, n, 8, nt=nt, fp64=fp64, seed=130 + k, with_obs=False)
            d, s = _shot_dev(capi, sh)
            tr = capi.DeviceArray(sh["nt"] * len(sh["rec"]) * np.dtype(sh["dtype"]).itemsize)
            shots.append((sh, d, s, tr))
        refs = [oracle.forward(sh["o"], sh["coef"], sh["iv"], sh["nt"], sh["src"], sh["wavelet"], sh["rec"])
                for sh, _, _, _ in shots]
        b = capi.Batch([capi.hsx_grid(x[0]["p"]) for x in shots], [x[1]["coef"] for x in shots],
                       [x[2] for x in shots], [x[3] for x in shots])
        shapes = [(32, 8, -1, 0), (64, 8, -1, 0)] if fp64 else [(64, 16, -1, 1), (32, 8, -1, 0)]
        for shape in shapes:
            print(f"persistent batch fp64={fp64} shots={len(sizes)} shape {shape}", flush=True)
            b.set_shape(*shape)
            for use_graph in (True, False):
                for sh, _, _, tr in shots:        # stale traces would show as NaN
                    tr.upload(np.full(tr.nbytes // np.dtype(sh["dtype"]).itemsize, np.nan, sh["dtype"]))
                b.run(use_graph=use_graph)
                for (sh, d, s, tr), ref in zip(shots, refs):
                    _same(tr.download(sh["dtype"]).reshape(ref.shape), ref)
        with pytest.raises(capi.HsxError):
            b.set_shape(48, 8, -1, 0)
        b.close()


def test_plan_with_autotune_bit_exact(oracle, capi):
    """a plan made with HSX_ADJ_AUTOTUNE (the layout's tile shape tuned first)
    gives the oracle's gradient bit for bit"""
    sh = make_shot(oracle, (40, 36, 32), 8, nt=30, seed=37)
    g = capi.hsx_grid(sh["p"])
    d, s = _shot_dev(capi, sh)
    capi.autotune_clear()
    try:
        plan = capi.AdjointPlan(g, sh["nt"], len(sh["rec"]), 6, capi.HSX_CKPT_FP32, autotune=True)
        grad = _dev(capi, np.zeros(sh["p"].npoint))
        plan.gradient(g, d["coef"], s, _dev(capi, sh["d_obs"]), grad)
        plan.close()
        gref, _ = oracle.gradient(sh["o"], sh["coef"], sh["iv"], sh["dt"], sh["nt"], sh["src"], sh["wavelet"],
                                  sh["rec"], sh["d_obs"])
        _same(grad.download(np.float64), gref)
    finally:
        capi.autotune_clear()
