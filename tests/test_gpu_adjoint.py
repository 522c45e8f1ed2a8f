"""Discrete adjoint of LLF + WENO-Z + RK3 (periodic), through the C ABI and
its torch.autograd.Function:

* the forward pass of the adjoint object is the plain solver: bit for bit
  the Solver's, and the PyTorch FP64 reference's (itself bitwise equal to the
  C oracle) to the workloads' tolerances;
* on a 32x32 instance the gradient equals the autograd gradient of the
  PyTorch reference (same frozen step sizes) to 1e-10 relative;
* the gradient does not depend on the checkpoint interval (bit for bit);
* on the 256x256 vortex with 64 Fourier modes, 200 steps at CFL 0.4 and a
  checkpoint every 20 steps, the gradient with respect to the mode
  amplitudes agrees with central finite differences (h = 1e-6, same frozen
  step sizes) to 1e-6 relative (l2 over the 64 amplitudes)."""
import importlib

import numpy as np
import pytest
import torch

from conftest import assert_dts, assert_parity, dense

import torch_ref

pytestmark = pytest.mark.gpu


def _adj():
    return importlib.import_module("2dense_amd.adjoint")


def _problem(n, n_modes, device):
    A = _adj()
    ic = A.FourierIC(n, n_modes=n_modes, device=device)
    theta = torch.tensor(A.draw_amplitudes(n_modes), dtype=torch.float64, device=device)
    target = ic(torch.tensor(A.draw_amplitudes(n_modes, seed=12), dtype=torch.float64, device=device))[0]
    return ic, theta, target


def test_adjoint_forward_is_the_solver():
    A = _adj()
    n, steps = 32, 20
    ic, theta, _ = _problem(n, 64, "cpu")
    U0 = ic(theta)
    ref, ref_dts = torch_ref.run(U0, steps, cfl=0.4, dx=ic.h, dy=ic.h)
    print("adjoint forward 32x32, 20 steps", flush=True)
    R = A.Rollout(n, n, steps, dx=ic.h, dy=ic.h, cfl=0.4, checkpoint_every=6)
    uN, dts = R.forward(U0.cuda())
    assert_dts(dts, ref_dts)
    assert_parity(uN.cpu().numpy(), ref.numpy(), ic.h, ic.h, "adjoint forward")
    # the plain solver from the same state
    D = dense()
    cfg = D.make_cfg(n, n, benchmark=0, splitting=D.LLF, recon=D.WENOZ, cfl=0.4, dx=ic.h, dy=ic.h,
                     bc=[D.BC_PERIODIC] * 4)
    with D.Solver(cfg) as s:
        s.set_state(U0.numpy())
        taken, sdts = s.run(steps, want_dts=True)
        assert np.array_equal(sdts, dts) and np.array_equal(s.interior(), uN.cpu().numpy())
    # a given schedule is used as given
    uN2, dts2 = R.forward(U0.cuda(), dt_schedule=dts)
    assert np.array_equal(dts2, dts) and torch.equal(uN2, uN)


@pytest.mark.parametrize("steps,k", [(20, 6), (7, 7), (5, 1)])
def test_adjoint_vs_torch_reference_gradient(steps, k):
    A = _adj()
    n = 32
    ic, theta, target = _problem(n, 64, "cpu")
    # reference: autograd through the PyTorch restatement with the same frozen steps
    _, dts = torch_ref.run(ic(theta), steps, cfl=0.4, dx=ic.h, dy=ic.h)
    th = theta.clone().requires_grad_(True)
    U0 = ic(th)
    uN_ref, _ = torch_ref.run(U0, steps, dx=ic.h, dy=ic.h, dts=dts)
    J_ref = A.density_mismatch(uN_ref, target, ic.h, ic.h)
    g_ref_th, g_ref_u0 = torch.autograd.grad(J_ref, [th, U0])

    print(f"adjoint gradient 32x32, {steps} steps, checkpoint every {k}", flush=True)
    R = A.Rollout(n, n, steps, dx=ic.h, dy=ic.h, cfl=0.4, checkpoint_every=k)
    # gradient with respect to the conserved initial state (same U0 as the reference)
    U0d = U0.detach().cuda().requires_grad_(True)
    uN = A.rollout(U0d, R)
    assert_dts(R.dts, dts)
    J = A.density_mismatch(uN, target.cuda(), ic.h, ic.h)
    J.backward()
    assert J.detach().item() == pytest.approx(J_ref.detach().item(), rel=1e-11)
    gu = U0d.grad.cpu().numpy()
    gu_ref = g_ref_u0.numpy()
    err_u = np.abs(gu - gu_ref).max() / np.abs(gu_ref).max()
    # gradient with respect to the mode amplitudes, through the GPU parameterisation
    icd, thetad, targetd = _problem(n, 64, "cuda")
    thd = thetad.clone().requires_grad_(True)
    J2 = A.density_mismatch(A.rollout(icd(thd), R, dts), targetd, ic.h, ic.h)
    J2.backward()
    err_th = np.linalg.norm(thd.grad.cpu().numpy() - g_ref_th.numpy()) / np.linalg.norm(g_ref_th.numpy())
    print(f"dJ/dU0 max rel {err_u:.3e}, dJ/dtheta l2 rel {err_th:.3e}", flush=True)
    assert err_u < 1e-10
    assert err_th < 1e-10


def test_adjoint_checkpoint_interval_invariance():
    A = _adj()
    n, steps = 48, 30
    ic, theta, target = _problem(n, 16, "cuda")
    U0 = ic(theta)
    g = torch.randn(4, n, n, dtype=torch.float64, generator=torch.Generator().manual_seed(3)).cuda()
    outs = []
    for k in (1, 7, 30):
        print(f"checkpoint every {k}", flush=True)
        R = A.Rollout(n, n, steps, dx=ic.h, dy=ic.h, cfl=0.4, checkpoint_every=k)
        R.forward(U0)
        outs.append(R.backward(g))
    assert torch.equal(outs[0], outs[1]) and torch.equal(outs[0], outs[2])


def test_adjoint_stale_graph_raises():
    A = _adj()
    n = 16
    ic, theta, target = _problem(n, 4, "cuda")
    R = A.Rollout(n, n, 3, dx=ic.h, dy=ic.h, cfl=0.4, checkpoint_every=2)
    U0 = ic(theta).requires_grad_(True)
    uN = A.rollout(U0, R)
    A.rollout(U0.detach(), R)  # second forward pass replaces the checkpoints
    with pytest.raises(RuntimeError):
        uN.sum().backward()


def test_adjoint_rejects_unsupported():
    D = dense()
    A = _adj()
    cfg = D.make_cfg(32, 32, benchmark=0, splitting=D.CH, recon=D.WENOZ, dx=0.1, dy=0.1, bc=[D.BC_PERIODIC] * 4)
    import ctypes as C
    lib = A._bind(D.load_library())
    h = C.c_void_p()
    assert lib.dense_adj_create(C.byref(cfg), 10, 5, C.byref(h)) == -2
    cfg = D.make_cfg(32, 32, benchmark=0, splitting=D.LLF, recon=D.WENOZ, dx=0.1, dy=0.1,
                     bc=[D.BC_EXTRAP] * 2 + [D.BC_PERIODIC] * 2)
    assert lib.dense_adj_create(C.byref(cfg), 10, 5, C.byref(h)) == -2


def test_adjoint_fourier_gradient_vs_central_differences_256():
    """The differentiation workload: 256x256 periodic vortex, density modulated
    by 64 Fourier modes (amplitudes N(0, 1e-3^2)), J = sum (rho_200 - rho*)^2 dx dy."""
    A = _adj()
    n, steps, k, h = 256, 200, 20, 1e-6
    ic, theta, target = _problem(n, 64, "cuda")
    R = A.Rollout(n, n, steps, dx=ic.h, dy=ic.h, cfl=0.4, checkpoint_every=k)
    th = theta.clone().requires_grad_(True)
    print("adjoint 256x256, 200 steps, checkpoint every 20", flush=True)
    J = A.density_mismatch(A.rollout(ic(th), R), target, ic.h, ic.h)
    J.backward()
    g = th.grad.cpu().numpy()
    dts = R.dts.copy()
    fwd_ms, bwd_ms = R.last_ms()
    print(f"J = {J.detach().item():.6e}; forward {fwd_ms:.2f} ms, backward {bwd_ms:.2f} ms", flush=True)

    def Jat(t):
        uN, _ = R.forward(ic(t), dt_schedule=dts)
        return float(A.density_mismatch(uN, target, ic.h, ic.h))

    fd = np.zeros(64)
    for m in range(64):
        e = torch.zeros_like(theta)
        e[m] = h
        fd[m] = (Jat(theta + e) - Jat(theta - e)) / (2 * h)
    err = np.linalg.norm(g - fd) / np.linalg.norm(fd)
    worst = np.max(np.abs(g - fd) / np.abs(fd))
    print(f"dJ/dtheta vs central differences: l2 rel {err:.3e}, worst mode rel {worst:.3e}", flush=True)
    assert err < 1e-6
