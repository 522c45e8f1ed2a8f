"""Gradients through the explicit update: the hand-written FP64 discrete adjoint
of Lax-Friedrichs splitting + WENO-Z + SSP-RK3 on a periodic block, behind a
``torch.autograd.Function``.

``Rollout`` binds the dense_adj_* part of the C ABI.  A forward pass advances a
device state ``(4, nj, ni)`` float64 by ``nsteps`` steps and keeps checkpoints
of U^n every ``checkpoint_every`` steps; the backward pass recomputes each
segment's stage states with the forward kernels and applies the transposed
step (vector-Jacobian products of the WENO-Z stencils and of the flux split).
The step sizes of the forward pass (CFL steps, time_advance_general.f90:20-24,
or a given schedule) are frozen, so the gradient is that of the discrete map
with those step sizes.

``rollout(u0, R)`` is the differentiable call.  ``FourierIC`` maps amplitudes
of seeded Fourier modes of the initial density to the conserved initial state
(differentiable torch code), and ``density_mismatch`` is the objective.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np
import torch

from . import Cfg, DenseError, LLF, WENOZ, BC_PERIODIC, load_library, make_cfg


def _bind(lib):
    if getattr(lib, "_adj_bound", False):
        return lib
    H = C.c_void_p
    dp = C.POINTER(C.c_double)
    vp = C.c_void_p
    lib.dense_adj_create.argtypes = [C.POINTER(Cfg), C.c_int, C.c_int, C.POINTER(H)]
    lib.dense_adj_destroy.argtypes = [H]
    lib.dense_adj_forward.argtypes = [H, vp, vp, dp, dp]
    lib.dense_adj_backward.argtypes = [H, vp, vp]
    lib.dense_adj_last_ms.argtypes = [H, dp, dp]
    lib.dense_adj_last_error.argtypes = [H]
    lib.dense_adj_last_error.restype = C.c_char_p
    for n in ("dense_adj_create", "dense_adj_destroy", "dense_adj_forward", "dense_adj_backward", "dense_adj_last_ms"):
        getattr(lib, n).restype = C.c_int
    lib._adj_bound = True
    return lib


def _dev_f64(x: torch.Tensor, shape) -> torch.Tensor:
    if not (isinstance(x, torch.Tensor) and x.is_cuda and x.dtype == torch.float64):
        raise TypeError("expected a float64 CUDA tensor")
    if tuple(x.shape) != tuple(shape):
        raise ValueError(f"expected shape {tuple(shape)}, got {tuple(x.shape)}")
    return x.contiguous()


class Rollout:
    """nsteps steps of LLF (101) + WENO-Z (255) + RK3 on a periodic ni x nj
    block with spacings dx, dy; differentiable through ``rollout``.

    A Rollout keeps the checkpoints of its last forward pass only: the
    backward pass of an earlier forward pass raises."""

    def __init__(self, ni: int, nj: int, nsteps: int, *, dx: float, dy: float, gamma: float = 1.4,
                 cfl: float = 0.4, checkpoint_every: int = 20, tile: int = 0):
        self.lib = _bind(load_library())
        self.cfg = make_cfg(ni, nj, benchmark=0, splitting=LLF, recon=WENOZ, cfl=cfl, gamma=gamma, dx=dx, dy=dy,
                            bc=[BC_PERIODIC] * 4, tile=tile)
        self.ni, self.nj, self.nsteps, self.k = ni, nj, nsteps, checkpoint_every
        h = C.c_void_p()
        self._check(self.lib.dense_adj_create(C.byref(self.cfg), nsteps, checkpoint_every, C.byref(h)), None)
        self.h = h
        self.generation = 0
        self.dts = None

    def _check(self, rc, h):
        if rc != 0:
            msg = self.lib.dense_adj_last_error(h)
            raise DenseError(rc, msg.decode() if msg else "")

    def close(self):
        if getattr(self, "h", None):
            self.lib.dense_adj_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def shape(self):
        return (4, self.nj, self.ni)

    def forward(self, u0: torch.Tensor, dt_schedule=None):
        """Returns (u_N, dts).  dt_schedule: None (CFL steps) or nsteps step sizes."""
        u0 = _dev_f64(u0, self.shape)
        out = torch.empty_like(u0)
        dts = np.zeros(self.nsteps)
        sched = None
        if dt_schedule is not None:
            sched = np.ascontiguousarray(np.asarray(dt_schedule, dtype=np.float64).reshape(-1))
            if sched.size != self.nsteps:
                raise ValueError(f"dt_schedule needs {self.nsteps} entries")
        torch.cuda.synchronize(u0.device)
        dp = C.POINTER(C.c_double)
        self._check(self.lib.dense_adj_forward(self.h, C.c_void_p(u0.data_ptr()), C.c_void_p(out.data_ptr()),
                                               sched.ctypes.data_as(dp) if sched is not None else None,
                                               dts.ctypes.data_as(dp)), self.h)
        self.generation += 1
        self.dts = dts
        return out, dts

    def backward(self, gN: torch.Tensor) -> torch.Tensor:
        gN = _dev_f64(gN, self.shape)
        g0 = torch.empty_like(gN)
        torch.cuda.synchronize(gN.device)
        self._check(self.lib.dense_adj_backward(self.h, C.c_void_p(gN.data_ptr()), C.c_void_p(g0.data_ptr())),
                    self.h)
        return g0

    def last_ms(self):
        f, b = C.c_double(), C.c_double()
        self._check(self.lib.dense_adj_last_ms(self.h, C.byref(f), C.byref(b)), self.h)
        return f.value, b.value


class _RolloutFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, u0, R, dt_schedule):
        out, _ = R.forward(u0, dt_schedule)
        ctx.R = R
        ctx.gen = R.generation
        return out

    @staticmethod
    def backward(ctx, gN):
        R = ctx.R
        if R.generation != ctx.gen:
            raise RuntimeError("Rollout was run forward again after this graph was built; its checkpoints are gone")
        return R.backward(gN), None, None


def rollout(u0: torch.Tensor, R: Rollout, dt_schedule=None) -> torch.Tensor:
    """u_N = nsteps steps from u0 (4, nj, ni) float64 CUDA; differentiable in u0."""
    return _RolloutFn.apply(u0, R, dt_schedule)


# ------------------------------------------------------------------ problem

def vortex_prim(n: int, L: float = 10.0, eps: float = 5.0, xc: float = 5.0, yc: float = 5.0,
                gamma: float = 1.4):
    """Isentropic vortex of strength eps at (xc, yc) in the free stream
    rho = u = v = p = 1 on the periodic square [0, L]^2 (cell centres,
    minimum-image distance).  Returns rho, u, v, p (n x n) and the spacing."""
    h = L / n
    x = (np.arange(n) + 0.5) * h
    X, Y = np.meshgrid(x, x)
    dxp = (X - xc + 0.5 * L) % L - 0.5 * L
    dyp = (Y - yc + 0.5 * L) % L - 0.5 * L
    r2 = dxp * dxp + dyp * dyp
    f = eps / (2.0 * math.pi) * np.exp(0.5 * (1.0 - r2))
    u = 1.0 - f * dyp
    v = 1.0 + f * dxp
    T = 1.0 - (gamma - 1.0) * eps * eps * np.exp(1.0 - r2) / (8.0 * gamma * math.pi ** 2)
    rho = T ** (1.0 / (gamma - 1.0))
    p = rho ** gamma
    return rho, u, v, p, h


class FourierIC:
    """U0(theta): the base vortex with its density multiplied by
    (1 + sum_m theta_m phi_m), phi_m = cos(2 pi (kx_m x + ky_m y)/L + phase_m),
    velocity and pressure kept.  Wavenumbers kx, ky are drawn from 0..kmax
    (not both zero) and phases from [0, 2 pi) with the given seed."""

    def __init__(self, n: int, n_modes: int = 64, seed: int = 7, kmax: int = 6, L: float = 10.0,
                 gamma: float = 1.4, eps: float = 5.0, device="cpu"):
        rng = np.random.default_rng(seed)
        rho, u, v, p, h = vortex_prim(n, L=L, eps=eps, gamma=gamma)
        self.n, self.h, self.gamma, self.L = n, h, gamma, L
        k = []
        while len(k) < n_modes:
            kx, ky = rng.integers(0, kmax + 1, size=2)
            if kx == 0 and ky == 0:
                continue
            k.append((kx, ky))
        phase = rng.uniform(0.0, 2.0 * math.pi, size=n_modes)
        x = (np.arange(n) + 0.5) * h
        X, Y = np.meshgrid(x, x)
        basis = np.stack([np.cos(2.0 * math.pi * (kx * X + ky * Y) / L + ph) for (kx, ky), ph in zip(k, phase)])
        t = lambda a: torch.as_tensor(a, dtype=torch.float64, device=device)
        self.rho, self.u, self.v, self.p = t(rho), t(u), t(v), t(p)
        self.basis = t(basis)                     # (M, n, n)
        self.modes = k

    def __call__(self, theta: torch.Tensor) -> torch.Tensor:
        # elementwise product and sum (a tensordot backward becomes a slow library GEMM)
        mod = 1.0 + (theta[:, None, None] * self.basis).sum(0)
        rho = self.rho * mod
        E = self.p / (self.gamma - 1.0) + 0.5 * rho * (self.u * self.u + self.v * self.v)
        return torch.stack([rho, rho * self.u, rho * self.v, E])


def draw_amplitudes(n_modes: int = 64, sigma: float = 1e-3, seed: int = 11) -> np.ndarray:
    return np.random.default_rng(seed).normal(0.0, sigma, size=n_modes)


def density_mismatch(uN: torch.Tensor, rho_target: torch.Tensor, dx: float, dy: float) -> torch.Tensor:
    """J = sum over cells of (rho_N - rho*)^2 dx dy."""
    d = uN[0] - rho_target
    return (d * d).sum() * (dx * dy)
