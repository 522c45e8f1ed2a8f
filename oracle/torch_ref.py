"""Plain PyTorch FP64 restatement of LLF + WENO-Z + RK3 on a periodic block --
TEST INFRASTRUCTURE ONLY (the reference gradient of the adjoint checks).

Only tests/ and bench.py's checks import this module.  It restates, with torch
tensor operations in the order of the C oracle (dense_oracle.c), the pieces of
the reference's explicit update that the adjoint differentiates:

* primitive recovery, get_primary.f90:52-61 (pressure floor 1e-6 as a
  single-precision literal widened);
* Lax-Friedrichs split per cell, alpha = |u_n| + c, flux_splitting_LLF.f90:37-47
  (x) and :62-72 (y);
* WENO-Z, weno_mod.f90:50-95,168-188, on the stencils
  lf(m) = F+(f-3+m) and lf(m) = F-(f+2-m) of face f (between cells f-1 and f),
  scheme_weno5z_general.f90:67-166; h = h+ + h-;
* SSP-RK3 stages, time_advance_general.f90:35-217, and the CFL step
  time_advance_general.f90:20-24.

The boundary is periodic (torch.roll).  torch's autograd of this code is the
"plain PyTorch reference gradient"; its forward pass is pinned to the C oracle
by the CPU tests.
"""
from __future__ import annotations

import numpy as np
import torch

P_FLOOR = float(np.float32(1.0e-6))
WENO_SS = float(np.float32(1.0e-6))
A300, A301, A302 = 1.0 / 3.0, -7.0 / 6.0, 11.0 / 6.0
A310, A311, A312 = -1.0 / 6.0, 5.0 / 6.0, 1.0 / 3.0
A320, A321, A322 = 1.0 / 3.0, 5.0 / 6.0, -1.0 / 6.0
C30, C31, C32 = 0.1, 0.6, 0.3


class _Sqrt(torch.autograd.Function):
    # torch's vectorised CPU sqrt is not always the correctly rounded square
    # root that the oracle (and the GPU) take; numpy's is
    @staticmethod
    def forward(ctx, x):
        y = torch.from_numpy(np.sqrt(x.detach().cpu().numpy())).to(x.device)
        ctx.save_for_backward(y)
        return y

    @staticmethod
    def backward(ctx, g):
        (y,) = ctx.saved_tensors
        return g * 0.5 / y


def prim(U, gamma):
    rho = U[0]
    u = U[1] / rho
    v = U[2] / rho
    p = (gamma - 1.0) * (U[3] - 0.5 * rho * (u * u + v * v))
    p = torch.where(p < 0.0, torch.full_like(p, P_FLOOR), p)
    c = _Sqrt.apply(gamma * p / rho)
    return rho, u, v, p, c


def _div(x, c: float):
    # torch turns division by a Python scalar into multiplication by its
    # reciprocal; divide by a tensor to keep the IEEE quotient of the oracle
    return x / torch.full_like(x, c)


def weno5z(l0, l1, l2, l3, l4):
    q0 = A300 * l0 + A301 * l1 + A302 * l2
    q1 = A310 * l1 + A311 * l2 + A312 * l3
    q2 = A320 * l2 + A321 * l3 + A322 * l4
    x = (l0 - 2.0 * l1) + l2
    y = (l0 - 4.0 * l1) + 3.0 * l2
    is0 = _div(13.0 * (x * x), 12.0) + _div(y * y, 4.0)
    x = (l1 - 2.0 * l2) + l3
    y = l1 - l3
    is1 = _div(13.0 * (x * x), 12.0) + _div(y * y, 4.0)
    x = (l2 - 2.0 * l3) + l4
    y = (3.0 * l2 - 4.0 * l3) + l4
    is2 = _div(13.0 * (x * x), 12.0) + _div(y * y, 4.0)
    tau5 = torch.abs(is2 - is0)
    t = tau5 / (WENO_SS + is0)
    a0 = C30 * (1.0 + t * t)
    t = tau5 / (WENO_SS + is1)
    a1 = C31 * (1.0 + t * t)
    t = tau5 / (WENO_SS + is2)
    a2 = C32 * (1.0 + t * t)
    s = a0 + a1 + a2
    w0, w1, w2 = a0 / s, a1 / s, a2 / s
    return w0 * q0 + w1 * q1 + w2 * q2


def face_flux(U, gamma, axis):
    """h(f) for every face f along `axis` (-1: x, -2: y), shape (4, nj, ni);
    h(f) is the face between cells f-1 and f."""
    rho, u, v, p, c = prim(U, gamma)
    un = u if axis == -1 else v
    a = torch.abs(un) + c
    upa, uma = un + a, un - a
    nrm = 1 if axis == -1 else 2
    out = []
    for k in range(4):
        if k == nrm:
            fp, fm = 0.5 * (upa * U[k] + p), 0.5 * (uma * U[k] + p)
        elif k == 3:
            fp, fm = 0.5 * (upa * U[k] + un * p), 0.5 * (uma * U[k] + un * p)
        else:
            fp, fm = 0.5 * (upa * U[k]), 0.5 * (uma * U[k])
        lp = [torch.roll(fp, 3 - m, dims=axis) for m in range(5)]
        lm = [torch.roll(fm, m - 2, dims=axis) for m in range(5)]
        out.append(weno5z(*lp) + weno5z(*lm))
    return torch.stack(out)


def cfl_dt(U, gamma, cfl, dx, dy) -> float:
    with torch.no_grad():
        _, u, v, _, c = prim(U, gamma)
        ai = float((torch.abs(u) + c).max())
        aj = float((torch.abs(v) + c).max())
    dt_xi, dt_eta = dx / ai, dy / aj
    return cfl * dt_xi * dt_eta / (dt_xi + dt_eta)


def step(U0, dt, gamma, dx, dy):
    """One SSP-RK3 step, each stage written as time_advance_general.f90 writes it."""
    def diffs(U):
        hx = face_flux(U, gamma, -1)
        he = face_flux(U, gamma, -2)
        return torch.roll(hx, -1, dims=-1) - hx, torch.roll(he, -1, dims=-2) - he

    dxi, deta = diffs(U0)
    U1 = U0 - dt / dx * dxi - dt / dy * deta
    dxi, deta = diffs(U1)
    U2 = 3.0 / 4.0 * U0 + 1.0 / 4.0 * U1 - 1.0 / 4.0 * (dt / dx * dxi + dt / dy * deta)
    dxi, deta = diffs(U2)
    return 1.0 / 3.0 * U0 + 2.0 / 3.0 * U2 - 2.0 / 3.0 * (dt / dx * dxi + dt / dy * deta)


def run(U0, nsteps, *, gamma=1.4, cfl=0.4, dx, dy, dts=None):
    """nsteps steps; dts None takes the CFL step of each step (from the
    detached state, so the schedule is not differentiated).  Returns (U_N, dts)."""
    U = U0
    used = []
    for n in range(nsteps):
        dt = float(dts[n]) if dts is not None else cfl_dt(U, gamma, cfl, dx, dy)
        used.append(dt)
        U = step(U, dt, gamma, dx, dy)
    return U, np.array(used)
