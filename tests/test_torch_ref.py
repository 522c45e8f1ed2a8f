"""The plain PyTorch FP64 reference of LLF + WENO-Z + RK3 (oracle/torch_ref.py)
that the adjoint's gradient is checked against: its forward pass equals the C
oracle bit for bit, and its autograd gradient agrees with central finite
differences of itself."""
import importlib

import numpy as np
import pytest
import torch

from conftest import dense  # noqa: F401  (puts the repository root and oracle/ on sys.path)

import pyoracle
import torch_ref


def _adj():
    return importlib.import_module("2dense_amd.adjoint")


def _oracle_run(U0, n, steps, h, dts=None):
    p = pyoracle.problem_params(1, n, n, splitting=101, recon=255, cfl=0.4)
    p.dx = p.dy = h
    for s in range(4):
        p.bc[s] = pyoracle.BC_PERIODIC
    o = pyoracle.Oracle(p)
    pad = np.zeros(o.shape_cons)
    pad[:, 4:-4, 4:-4] = U0
    o.set_cons(pad)
    used = []
    for k in range(steps):
        o.step(1)
        used.append(o.dt)
    return o.interior(), np.array(used)


def test_torch_ref_forward_is_the_oracle():
    A = _adj()
    n, steps = 32, 12
    ic = A.FourierIC(n)
    U0 = ic(torch.tensor(A.draw_amplitudes(), dtype=torch.float64))
    ref, ref_dts = _oracle_run(U0.numpy(), n, steps, ic.h)
    got, dts = torch_ref.run(U0, steps, cfl=0.4, dx=ic.h, dy=ic.h)
    assert np.array_equal(dts, ref_dts)
    assert np.array_equal(got.numpy(), ref)


def test_torch_ref_gradient_vs_central_differences():
    A = _adj()
    n, steps, h = 32, 6, 1e-6
    ic = A.FourierIC(n, n_modes=8)
    theta = torch.tensor(A.draw_amplitudes(8), dtype=torch.float64)
    target = ic(torch.tensor(A.draw_amplitudes(8, seed=12), dtype=torch.float64))[0]
    _, dts = torch_ref.run(ic(theta), steps, cfl=0.4, dx=ic.h, dy=ic.h)

    def J(th):
        uN, _ = torch_ref.run(ic(th), steps, dx=ic.h, dy=ic.h, dts=dts)
        return A.density_mismatch(uN, target, ic.h, ic.h)

    th = theta.clone().requires_grad_(True)
    g = torch.autograd.grad(J(th), th)[0].numpy()
    fd = np.zeros_like(g)
    for m in range(len(g)):
        e = torch.zeros_like(theta)
        e[m] = h
        fd[m] = (float(J(theta + e)) - float(J(theta - e))) / (2 * h)
    err = np.linalg.norm(g - fd) / np.linalg.norm(fd)
    print("relative l2 error of the torch gradient vs central differences", err)
    assert err < 1e-6
