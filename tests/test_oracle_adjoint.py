"""The CPU oracle of shot modeling and of the misfit gradient, pinned by central
finite differences of the misfit and by the linearity of the modeling in the
source (no reference counterpart: the reference's adjoint is a stub,
src/data_Acquisition.cpp:307-310)."""
import numpy as np
import pytest

from tests.shots import make_shot


def _misfit(oracle, p, o, coef, iv, shot, d_obs):
    tr = oracle.forward(o, coef, iv, shot["nt"], shot["src"], shot["wavelet"], shot["rec"])
    r = (tr - d_obs).astype(np.float64)
    return 0.5 * float(np.sum(r * r))


@pytest.mark.parametrize("order", [4, 8])
def test_gradient_matches_central_differences(oracle, order):
    sh = make_shot(oracle, (18, 16, 14), order, nt=70, fp64=True, seed=order)
    o, iv, dt = sh["o"], sh["iv"], sh["dt"]
    g, J = oracle.gradient(o, sh["coef"], iv, dt, sh["nt"], sh["src"], sh["wavelet"], sh["rec"], sh["d_obs"])
    assert J > 0
    assert J == pytest.approx(_misfit(oracle, sh["p"], o, sh["coef"], iv, sh, sh["d_obs"]), rel=1e-12)
    rng = np.random.default_rng(1)
    inner = sh["inner_mask"]
    for trial in range(2):
        dm = np.where(inner, rng.standard_normal(o.npoint), 0.0) * sh["m"] * 0.05
        eps = 1e-4
        mp, mm = sh["m"] + eps * dm, sh["m"] - eps * dm
        cp = np.where(inner, dt * dt / mp, 0.0)
        cm = np.where(inner, dt * dt / mm, 0.0)
        Jp = _misfit(oracle, sh["p"], o, cp, iv, sh, sh["d_obs"])
        Jm = _misfit(oracle, sh["p"], o, cm, iv, sh, sh["d_obs"])
        fd = (Jp - Jm) / (2 * eps)
        adj = float(np.dot(g, dm))
        assert adj == pytest.approx(fd, rel=2e-6), (trial, adj, fd)


def test_gradient_is_zero_off_the_interior(oracle):
    """boundary planes and halos are not touched; grad is linear in the residual"""
    sh = make_shot(oracle, (14, 13, 12), 8, nt=40, fp64=True, seed=3)
    o, iv, dt = sh["o"], sh["iv"], sh["dt"]
    g, _ = oracle.gradient(o, sh["coef"], iv, dt, sh["nt"], sh["src"], sh["wavelet"], sh["rec"], sh["d_obs"])
    assert np.all(g[~sh["grad_mask"]] == 0)
    assert np.any(g[sh["grad_mask"]] != 0)
    # J is quadratic in the residual: scaling (d - d_obs) by 2 scales the gradient by 2
    d = oracle.forward(o, sh["coef"], iv, sh["nt"], sh["src"], sh["wavelet"], sh["rec"])
    d2_obs = d - 2.0 * (d - sh["d_obs"])
    g2, _ = oracle.gradient(o, sh["coef"], iv, dt, sh["nt"], sh["src"], sh["wavelet"], sh["rec"], d2_obs)
    np.testing.assert_allclose(g2, 2.0 * g, rtol=1e-10, atol=1e-14 * np.abs(g).max())


def test_forward_is_linear_in_the_source(oracle):
    sh = make_shot(oracle, (14, 13, 12), 4, nt=50, fp64=True, seed=5)
    o, iv = sh["o"], sh["iv"]
    a = oracle.forward(o, sh["coef"], iv, sh["nt"], sh["src"], sh["wavelet"], sh["rec"])
    b = oracle.forward(o, sh["coef"], iv, sh["nt"], sh["src"], 3.0 * sh["wavelet"], sh["rec"])
    np.testing.assert_allclose(b, 3.0 * a, rtol=1e-12, atol=1e-15 * np.abs(a).max())
    assert np.abs(a).max() > 0


def test_fp32_gradient_close_to_fp64(oracle):
    sh64 = make_shot(oracle, (16, 15, 14), 8, nt=60, fp64=True, seed=7)
    sh32 = make_shot(oracle, (16, 15, 14), 8, nt=60, fp64=False, seed=7)
    g64, _ = oracle.gradient(sh64["o"], sh64["coef"], sh64["iv"], sh64["dt"], sh64["nt"], sh64["src"],
                             sh64["wavelet"], sh64["rec"], sh64["d_obs"])
    g32, _ = oracle.gradient(sh32["o"], sh32["coef"], sh32["iv"], sh32["dt"], sh32["nt"], sh32["src"],
                             sh32["wavelet"], sh32["rec"], sh32["d_obs"])
    err = np.linalg.norm(g32 - g64) / np.linalg.norm(g64)
    assert err < 1e-4, err


@pytest.mark.parametrize("fp64", [True, False])
def test_given_residual_equals_formed_residual(oracle, fp64):
    """the gradient of a residual passed in (the adjoint source of any loss)
    equals the one formed from the traces and the data, bit for bit"""
    sh = make_shot(oracle, (14, 13, 12), 8, nt=40, fp64=fp64, seed=13)
    o, iv, dt = sh["o"], sh["iv"], sh["dt"]
    args = (o, sh["coef"], iv, dt, sh["nt"], sh["src"], sh["wavelet"], sh["rec"])
    g, J = oracle.gradient(*args, sh["d_obs"])
    d = oracle.forward(o, sh["coef"], iv, sh["nt"], sh["src"], sh["wavelet"], sh["rec"])
    r = (d - sh["d_obs"]).astype(sh["dtype"])
    g_r, J_r = oracle.gradient(*args, r, residual=True)
    assert J_r == J
    assert np.array_equal(g_r, g)
    # and it is linear in that residual
    g_s, _ = oracle.gradient(*args, (r * 0.25).astype(sh["dtype"]), residual=True)
    np.testing.assert_allclose(g_s, 0.25 * g, rtol=1e-5 if not fp64 else 1e-12, atol=1e-12 * np.abs(g).max())


def test_source_on_a_boundary_plane(oracle):
    """a source on the n1 boundary plane: the gradient stays zero off the points
    the free surface does not touch and still matches central differences"""
    sh = make_shot(oracle, (16, 14, 12), 8, nt=50, fp64=True, seed=17)
    p, o, iv, dt = sh["p"], sh["o"], sh["iv"], sh["dt"]
    r = p.inner()
    mid2, mid3 = (r[2] + r[3]) // 2, (r[4] + r[5]) // 2
    sh["src"] = np.array([r[0] + mid2 * p.n1 + mid3 * p.n1 * p.n2, (r[0] + 3) + (mid2 + 1) * p.n1 + mid3 * p.n1 * p.n2],
                         np.int64)
    sh["wavelet"] = np.repeat(sh["wavelet"], 2, axis=1).copy()
    d_obs = oracle.forward(o, sh["coef_true"], iv, sh["nt"], sh["src"], sh["wavelet"], sh["rec"])
    g, J = oracle.gradient(o, sh["coef"], iv, dt, sh["nt"], sh["src"], sh["wavelet"], sh["rec"], d_obs)
    assert J > 0
    assert np.all(g[~sh["grad_mask"]] == 0)
    rng = np.random.default_rng(5)
    dm = np.where(sh["inner_mask"], rng.standard_normal(o.npoint), 0.0) * sh["m"] * 0.05
    eps = 1e-4
    Jp = _misfit(oracle, p, o, np.where(sh["inner_mask"], dt * dt / (sh["m"] + eps * dm), 0.0), iv, sh, d_obs)
    Jm = _misfit(oracle, p, o, np.where(sh["inner_mask"], dt * dt / (sh["m"] - eps * dm), 0.0), iv, sh, d_obs)
    assert float(np.dot(g, dm)) == pytest.approx((Jp - Jm) / (2 * eps), rel=2e-6)
