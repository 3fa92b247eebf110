"""The torch.autograd harness of the hand-written adjoint: loss.backward()
through hsx_gradient against the oracle gradient and against central
differences of the loss computed with the GPU forward."""
import numpy as np
import pytest

from tests.shots import make_shot

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_cuda():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.fail("no GPU visible to torch")
    return torch


@pytest.mark.parametrize("fp64", [True, False])
def test_backward_matches_oracle_gradient(oracle, capi, torch_cuda, fp64):
    torch = torch_cuda
    from hpcscan_amd.autograd import Ac2Shot, modeling

    sh = make_shot(oracle, (26, 22, 20), 8, nt=50, fp64=fp64, seed=31)
    shot = Ac2Shot(sh["p"], sh["dt"], sh["src"], sh["wavelet"], sh["rec"])
    dt = torch.float64 if fp64 else torch.float32
    m = torch.tensor(sh["m"], dtype=torch.float64, device="cuda", requires_grad=True)
    d_obs = torch.tensor(sh["d_obs"], dtype=dt, device="cuda")
    tr = modeling(m, shot)
    loss = 0.5 * ((tr.double() - d_obs.double()) ** 2).sum()
    loss.backward()
    gref, Jref = oracle.gradient(sh["o"], sh["coef"], sh["iv"], sh["dt"], sh["nt"], sh["src"], sh["wavelet"],
                                 sh["rec"], sh["d_obs"])
    got = m.grad.detach().cpu().numpy()
    err = np.linalg.norm(got - gref) / np.linalg.norm(gref)
    print(f"fp64={fp64} loss={loss.item():.6e} oracle J={Jref:.6e} rel err={err:.3e}", flush=True)
    assert loss.item() == pytest.approx(Jref, rel=1e-6)
    assert err < (1e-10 if fp64 else 1e-5)


def test_directional_derivative_matches_central_differences(oracle, capi, torch_cuda):
    torch = torch_cuda
    from hpcscan_amd.autograd import Ac2Shot, modeling

    sh = make_shot(oracle, (26, 22, 20), 4, nt=60, fp64=True, seed=41)
    shot = Ac2Shot(sh["p"], sh["dt"], sh["src"], sh["wavelet"], sh["rec"])
    d_obs = torch.tensor(sh["d_obs"], dtype=torch.float64, device="cuda")

    def loss_of(mt):
        return 0.5 * ((modeling(mt, shot) - d_obs) ** 2).sum()

    m0 = torch.tensor(sh["m"], dtype=torch.float64, device="cuda")
    m = m0.clone().requires_grad_(True)
    loss_of(m).backward()
    rng = np.random.default_rng(2)
    dm = torch.tensor(np.where(sh["inner_mask"], rng.standard_normal(m0.numel()), 0.0) * sh["m"] * 0.05,
                      device="cuda")
    eps = 1e-4
    with torch.no_grad():
        fd = (loss_of(m0 + eps * dm) - loss_of(m0 - eps * dm)).item() / (2 * eps)
    adj = float(torch.dot(m.grad, dm))
    print(f"adjoint {adj:.9e} central difference {fd:.9e}", flush=True)
    assert adj == pytest.approx(fd, rel=1e-6)


@pytest.mark.parametrize("scale,weighted", [(1e-6, False), (1.0, True), (3e-7, True)])
def test_backward_of_scaled_and_weighted_losses(oracle, capi, torch_cuda, scale, weighted):
    """losses other than the plain misfit: a tiny scale and per-sample weights.
    The trace gradient reaches the adjoint as the residual itself, so the
    gradient is the oracle's for that residual (FP32 grids)"""
    torch = torch_cuda
    from hpcscan_amd.autograd import Ac2Shot, modeling

    sh = make_shot(oracle, (26, 22, 20), 8, nt=50, fp64=False, seed=33)
    shot = Ac2Shot(sh["p"], sh["dt"], sh["src"], sh["wavelet"], sh["rec"])
    m = torch.tensor(sh["m"], dtype=torch.float64, device="cuda", requires_grad=True)
    d_obs = torch.tensor(sh["d_obs"], dtype=torch.float32, device="cuda")
    w = torch.tensor(np.random.default_rng(3).uniform(0.1, 2.0, sh["d_obs"].shape) if weighted
                     else np.ones(sh["d_obs"].shape), dtype=torch.float64, device="cuda")
    tr = modeling(m, shot)
    loss = scale * 0.5 * (w * (tr.double() - d_obs.double()) ** 2).sum()
    loss.backward()
    # the residual autograd hands to the adjoint: d loss / d traces, in FP32
    tr_h = tr.detach().double().cpu().numpy()
    res = (scale * w.cpu().numpy() * (tr_h - sh["d_obs"].astype(np.float64))).astype(np.float32)
    gref, _ = oracle.gradient(sh["o"], sh["coef"], sh["iv"], sh["dt"], sh["nt"], sh["src"], sh["wavelet"],
                              sh["rec"], res, residual=True)
    got = m.grad.detach().cpu().numpy()
    err = np.linalg.norm(got - gref) / np.linalg.norm(gref)
    print(f"scale={scale} weighted={weighted}: rel err {err:.3e}", flush=True)
    assert np.linalg.norm(gref) > 0
    assert err < 1e-5
