"""Discrete adjoint of the transient explicit step (fuss_adj_*) on the GPU.

The gradient of the probe misfit L = sum (T_probe - T_obs)^2 over every step
with respect to the segment fluxes q, the conductivity scale c and the
initial state is checked against central finite differences of the same GPU
forward model (h = 1e-6 relative, seeded directions) to 1e-6 relative; the
FP32-field variant against the FP64 adjoint; and a few gradient steps must
reduce the misfit of the inverse problem.
"""
import numpy as np
import pytest

from fuss_amd import cases
from tests.conftest import gpu_available

pytestmark = pytest.mark.gpu
FD_RTOL = 1e-6


@pytest.fixture(scope="module")
def adj():
    if not gpu_available():
        pytest.skip("no GPU")
    from fuss_amd import adjoint

    return adjoint


def make(adj, n_rk=3, precision=0, n=48, nseg=6, nsteps=45, ckpt=10, nprobe=8):
    case, info = cases.flux_plate(n=n, nseg=nseg, nprobe=nprobe, n_rk=n_rk, probe_rows=6)
    prob = adj.FluxProblem(case, info["seg_entries"], info["probes"], nsteps, ckpt, precision)
    return prob, info


def misfit(prob, q, c, obs):
    return float(((prob.forward(q, c) - obs) ** 2).sum())


def adjoint_grad(adj, prob, q0, c0, obs):
    import torch

    qt = torch.tensor(q0, dtype=torch.float64, requires_grad=True)
    ct = torch.tensor(c0, dtype=torch.float64, requires_grad=True)
    hist = adj.ProbeHistory.apply(qt, ct, prob)
    loss = ((hist - obs) ** 2).sum()
    loss.backward()
    return np.concatenate([qt.grad.numpy(), [ct.grad.item()]]), float(loss.detach())


@pytest.mark.parametrize("n_rk", [1, 3, 4])
def test_gradient_vs_central_differences(adj, n_rk):
    prob, info = make(adj, n_rk=n_rk)
    nseg = info["q_true"].size
    obs = prob.forward(info["q_true"], 1.0)
    q0, c0 = np.full(nseg, 3.0e4), 1.15
    g, loss = adjoint_grad(adj, prob, q0, c0, obs)
    assert loss > 0 and np.all(np.isfinite(g))
    rng = np.random.default_rng(100 + n_rk)
    scale = np.concatenate([np.full(nseg, 1.0e4), [1.0]])
    h = 1e-6
    for _ in range(3):
        v = rng.normal(size=nseg + 1) * scale
        Lp = misfit(prob, q0 + h * v[:-1], c0 + h * v[-1], obs)
        Lm = misfit(prob, q0 - h * v[:-1], c0 - h * v[-1], obs)
        fd = (Lp - Lm) / (2 * h)
        ad = float(g @ v)
        assert abs(fd - ad) <= FD_RTOL * abs(ad), (fd, ad)
    prob.close()


def test_initial_state_gradient(adj):
    """dL/dT^0.  The initial field is seeded non-uniform: from a uniform field
    every cell attains the global-min dt (a kink of min), and a perturbation
    of T^0 would pick a different minimising cell on either side."""
    import torch

    prob, info = make(adj, n_rk=3, nsteps=30)
    prob.T0[0] = 300.0 + 10.0 * np.random.default_rng(3).uniform(size=prob.T0[0].shape)
    obs = prob.forward(info["q_true"], 1.0)
    q0, c0 = np.full(info["q_true"].size, 2.5e4), 0.9
    hist = prob.forward(q0, c0)
    _, _, gT = prob.backward(2.0 * (hist - obs), grad_T0=True)
    T0 = prob.T0[0].copy()
    rng = np.random.default_rng(7)
    v = rng.uniform(0.0, 1.0, size=T0.shape)
    h = 1e-6

    def L_at(T):
        prob.T0[0] = T
        return misfit(prob, q0, c0, obs)

    fd = (L_at(T0 + h * v) - L_at(T0 - h * v)) / (2 * h)
    prob.T0[0] = T0
    ad = float(gT @ v.reshape(-1))
    assert abs(fd - ad) <= FD_RTOL * abs(ad), (fd, ad)
    assert torch.isfinite(torch.as_tensor(gT)).all()
    prob.close()


def test_fp32_adjoint_fields(adj):
    p64, info = make(adj, n_rk=3)
    p32, _ = make(adj, n_rk=3, precision=1)
    obs = p64.forward(info["q_true"], 1.0)
    q0, c0 = np.full(info["q_true"].size, 3.0e4), 1.1
    g64, _ = adjoint_grad(adj, p64, q0, c0, obs)
    g32, _ = adjoint_grad(adj, p32, q0, c0, obs)
    rel = np.linalg.norm(g32 - g64) / np.linalg.norm(g64)
    assert rel < 1e-4, rel
    p64.close()
    p32.close()


def test_stale_checkpoints_rejected(adj):
    import torch

    prob, info = make(adj, nsteps=12, ckpt=4)
    q = torch.tensor(np.full(info["q_true"].size, 2e4), requires_grad=True)
    c = torch.tensor(1.0, requires_grad=True)
    h1 = adj.ProbeHistory.apply(q, c, prob)
    adj.ProbeHistory.apply(q.detach(), c.detach(), prob)  # replaces the checkpoints
    with pytest.raises(RuntimeError):
        h1.sum().backward()
    prob.close()


def test_inverse_estimation_reduces_misfit(adj):
    """Gradient-based recovery of q and c from the probe histories."""
    import torch

    prob, info = make(adj, n_rk=3, nsteps=40)
    obs = prob.forward(info["q_true"], 1.0)
    nseg = info["q_true"].size
    x = torch.tensor(np.concatenate([np.full(nseg, 3.0), [1.2]]), dtype=torch.float64, requires_grad=True)
    opt = torch.optim.LBFGS([x], lr=1.0, max_iter=25, line_search_fn="strong_wolfe")

    def closure():
        opt.zero_grad()
        hist = adj.ProbeHistory.apply(x[:nseg] * 1e4, x[nseg], prob)
        loss = ((hist - obs) ** 2).sum()
        loss.backward()
        return loss

    L0 = float(closure().detach())
    opt.step(closure)
    L1 = float(closure().detach())
    assert L1 < 1e-3 * L0, (L0, L1)
    prob.close()
