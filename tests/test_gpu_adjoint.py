"""The GPU adjoint against the FP64 host adjoint, finite differences and PyTorch autograd.

Bar (the build's gradient target): 1e-4 relative agreement with the FP64 host adjoint."""
import numpy as np
import pytest

import oracle as orc
from lbm_amd import _abi as A
from lbm_amd import scenarios as sc
from lbm_amd.solver import Config, Solver

pytestmark = pytest.mark.gpu


def periodic_case(coll, nx=40, ny=24, seed=3):
    rng = np.random.default_rng(seed)
    om = 1.4
    S = (0.0, 1.1, 1.2, 0.0, 1.4, 0.0, 1.4, om, om) if coll == A.LBM_MRT else (0.0, om, om, 0.0, om, 0.0, om, om, om)
    cfg = Config(nx=nx, ny=ny, collision=coll, omega=om, S=S, periodic_x=True, periodic_y=True)
    rho = (1 + 0.01 * rng.uniform(-1, 1, (ny, nx))).astype(np.float32)
    u = (0.03 * rng.uniform(-1, 1, (ny, nx, 2))).astype(np.float32)
    F = (1e-4 * rng.uniform(-1, 1, (ny, nx, 2))).astype(np.float32)
    return cfg, rho, u, F


def rel(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


@pytest.mark.parametrize("coll", [A.LBM_BGK, A.LBM_MRT])
@pytest.mark.parametrize("T,K", [(30, 7), (24, 24), (12, 1)])
def test_gpu_adjoint_matches_fp64_host_adjoint(gpu_available, coll, T, K):
    print(f"case adjoint coll={coll} T={T} K={K}", flush=True)
    cfg, rho, u, F = periodic_case(coll)
    w = np.random.default_rng(5).normal(size=u.shape).astype(np.float32)
    s = Solver(cfg)
    s.set_force_field(F)
    s.init_from_macros(rho, u)
    s.adjoint_forward(T, K)
    gF, gU = s.adjoint_backward(w)
    o = orc.Oracle(cfg.to_c(), precision=64)
    o.set_force_field(F)
    o.init_from_macros(rho, u)
    rF, rU = o.adjoint(T, w)
    eF, eU = rel(gF, rF), rel(gU, rU)
    assert eF <= 1e-4, eF
    assert eU <= 1e-4, eU
    # the forward of the adjoint run is the ordinary step
    _, uT = s.get_macros()
    o2 = orc.Oracle(cfg.to_c(), precision=32)
    o2.set_force_field(F)
    o2.init_from_macros(rho, u)
    o2.step(T)
    _, u32 = o2.get_macros()
    assert float(np.max(np.abs(uT - u32))) < 1e-6
    s.close()


def test_gpu_adjoint_directional_finite_difference(gpu_available):
    cfg, rho, u, F = periodic_case(A.LBM_BGK, 64, 48)
    T = 40
    rng = np.random.default_rng(9)
    w = rng.normal(size=u.shape).astype(np.float32)
    s = Solver(cfg)
    s.set_force_field(F)
    s.init_from_macros(rho, u)
    s.adjoint_forward(T, 10)
    gF, _ = s.adjoint_backward(w)

    def loss(Fv):
        s2 = Solver(cfg)
        s2.set_force_field(Fv.astype(np.float32))
        s2.init_from_macros(rho, u)
        s2.step(T)
        _, uT = s2.get_macros()
        s2.close()
        return float(np.sum(w.astype(np.float64) * uT))

    for _ in range(4):
        d = rng.normal(size=F.shape)
        d *= 1e-3 / np.max(np.abs(d))  # FP32 forward: perturb well above its rounding
        fd = (loss(F + d) - loss(F - d)) / 2.0
        an = float(np.sum(gF * d))
        assert abs(fd - an) <= 1e-2 * abs(an), (fd, an)
    s.close()


def test_torch_autograd_function_and_inverse_problem(gpu_available):
    import torch

    from lbm_amd.autograd import Simulation, recover_force

    cfg, rho, u, _ = periodic_case(A.LBM_BGK, 48, 48)
    F_true = sc.fourier_force(48, 48, modes=4, amp=1e-5, seed=1)
    dev = "cuda"
    sim = Simulation(cfg, nsteps=60, checkpoint_every=20, rho0=rho)
    u0 = torch.as_tensor(u, device=dev)
    Ft = torch.as_tensor(F_true, device=dev, dtype=torch.float32)
    u_target = sim(Ft, u0).detach()
    # autograd gradient == the library's adjoint
    Fz = torch.zeros_like(Ft, requires_grad=True)
    uT = sim(Fz, u0)
    loss = torch.sum((uT - u_target) ** 2)
    loss.backward()
    s = Solver(cfg)
    s.set_force_field(np.zeros_like(F_true))
    s.init_from_macros(rho, u)
    s.adjoint_forward(60, 20)
    _, uT2 = s.get_macros()
    gF, _ = s.adjoint_backward(2 * (uT2 - u_target.cpu().numpy()))
    assert rel(Fz.grad.cpu().numpy(), gF) < 1e-5
    s.close()
    # Adam drives the misfit down
    _, losses = recover_force(sim, u_target, u0, iters=40, lr=2e-6)
    assert losses[-1] < 0.05 * losses[0], (losses[0], losses[-1])
    sim.close()


def test_autograd_backward_after_a_later_forward(gpu_available):
    """A second call of the same Simulation (here a no-grad validation forward with another
    force) before backward must not leak into the first node's gradient."""
    import torch

    from lbm_amd.autograd import Simulation
    from lbm_amd.solver import LbmError

    cfg, rho, u, _ = periodic_case(A.LBM_BGK, 32, 24)
    dev = "cuda"
    u0 = torch.as_tensor(u, device=dev)
    FA = torch.as_tensor(sc.fourier_force(32, 24, modes=3, amp=1e-5, seed=2), device=dev)
    FB = torch.as_tensor(sc.fourier_force(32, 24, modes=3, amp=3e-5, seed=5), device=dev)
    target = torch.zeros_like(u0)

    def grad_of(sim, F, between=None):
        Fv = F.clone().requires_grad_(True)
        loss = torch.sum((sim(Fv, u0) - target) ** 2)
        if between is not None:
            with torch.no_grad():
                sim(between, u0)
        loss.backward()
        return Fv.grad.cpu().numpy()

    sim = Simulation(cfg, nsteps=30, checkpoint_every=10, rho0=rho)
    g_clean = grad_of(sim, FA)
    g_after = grad_of(sim, FA, between=FB)
    g_b = grad_of(sim, FB)
    assert rel(g_after, g_clean) < 1e-6
    assert rel(g_b, g_clean) > 1e-3  # the two forces give different gradients
    # the C side refuses a backward whose forward state was changed underneath it
    s = sim.solver
    s.adjoint_forward(30, 10)
    s.set_force_field(FB)
    with pytest.raises(LbmError):
        s.adjoint_backward(np.zeros((24, 32, 2), np.float32))
    sim.close()
