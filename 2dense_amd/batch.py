"""Batched mode: an ensemble of independent blocks advanced together.

``Batch`` binds the dense_batch_* part of the C ABI: all instances advance by
one persistent launch per RK stage, each with its own grid, boundaries and
CFL step, and each evolves bit for bit as it would alone through ``Solver``.

``make_ensemble`` builds the seeded ensemble of isentropic vortices and
shock/vortex interactions used to measure the batched path; its initial
states are built either here (``vortex_state``, ``shock_vortex_state``) or on
the device from the instance parameters (``Batch.init_ic``).
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass

import numpy as np

from . import (BC_EXTRAP, BC_PERIODIC, BC_REFLECT, BF, Cfg, DenseError, LLF, WENOZ, _dp, load_library,
               make_cfg)


IC_VORTEX, IC_SHOCK_VORTEX, IC_SHOCK_VORTEX_REF = 1, 2, 3


class IcDesc(C.Structure):
    """dense_ic of include/dense_capi.h."""
    _fields_ = [("kind", C.c_int), ("eps", C.c_double), ("xc", C.c_double), ("yc", C.c_double),
                ("uinf", C.c_double), ("vinf", C.c_double), ("length", C.c_double), ("mach", C.c_double),
                ("xs", C.c_double), ("rc", C.c_double), ("alpha", C.c_double)]


def _bind(lib):
    if getattr(lib, "_batch_bound", False):
        return lib
    H = C.c_void_p
    dp = C.POINTER(C.c_double)
    ip = C.POINTER(C.c_int)
    lib.dense_batch_create.argtypes = [C.POINTER(Cfg), C.c_int, C.POINTER(H)]
    lib.dense_batch_destroy.argtypes = [H]
    lib.dense_batch_set_state.argtypes = [H, C.c_int, dp, C.c_double]
    lib.dense_batch_get_state.argtypes = [H, C.c_int, dp]
    lib.dense_batch_run.argtypes = [H, C.c_int, dp, dp]
    lib.dense_batch_set_schedule.argtypes = [H, C.c_int, C.c_int, C.c_int]
    lib.dense_batch_autotune.argtypes = [H, C.c_int, ip, ip, ip, dp]
    lib.dense_batch_last_run_gpu_ms.argtypes = [H, dp]
    lib.dense_batch_last_error.argtypes = [H]
    lib.dense_batch_last_error.restype = C.c_char_p
    lib.dense_batch_init_ic.argtypes = [H, C.POINTER(IcDesc)]
    for n in ("dense_batch_create", "dense_batch_destroy", "dense_batch_set_state", "dense_batch_get_state",
              "dense_batch_run", "dense_batch_set_schedule", "dense_batch_autotune", "dense_batch_last_run_gpu_ms",
              "dense_batch_init_ic"):
        getattr(lib, n).restype = C.c_int
    lib._batch_bound = True
    return lib


class Batch:
    """Ensemble of blocks sharing splitting/reconstruction/stages."""

    def __init__(self, cfgs):
        self.lib = _bind(load_library())
        self.cfgs = list(cfgs)
        arr = (Cfg * len(self.cfgs))(*self.cfgs)
        h = C.c_void_p()
        self._check(self.lib.dense_batch_create(arr, len(self.cfgs), C.byref(h)), None)
        self.h = h
        self.n = len(self.cfgs)
        self.t = np.zeros(self.n)

    def _check(self, rc, h):
        if rc != 0:
            msg = self.lib.dense_batch_last_error(h)
            raise DenseError(rc, msg.decode() if msg else "")

    def close(self):
        if getattr(self, "h", None):
            self.lib.dense_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _shape(self, i):
        c = self.cfgs[i]
        return (4, c.nj + 2 * BF, c.ni + 2 * BF)

    def set_state(self, i: int, cons: np.ndarray, t0: float = 0.0):
        shp = self._shape(i)
        a = np.asarray(cons, dtype=np.float64)
        if a.shape == (4, shp[1] - 2 * BF, shp[2] - 2 * BF):
            p = np.zeros(shp)
            p[:, BF:-BF, BF:-BF] = a
            a = p
        a = np.ascontiguousarray(a.reshape(shp))
        self._check(self.lib.dense_batch_set_state(self.h, i, _dp(a), t0), self.h)
        self.t[i] = t0

    def init_ic(self, insts):
        """Generate every instance's initial state on the device from its
        parameters (``Instance`` list as made by ``make_ensemble``, or
        ``IcDesc`` list); clocks restart at 0."""
        descs = [x if isinstance(x, IcDesc) else ic_desc(x) for x in insts]
        if len(descs) != self.n:
            raise ValueError(f"{len(descs)} descriptors for {self.n} instances")
        arr = (IcDesc * self.n)(*descs)
        self._check(self.lib.dense_batch_init_ic(self.h, arr), self.h)
        self.t = np.zeros(self.n)

    def interior(self, i: int) -> np.ndarray:
        a = np.empty(self._shape(i))
        self._check(self.lib.dense_batch_get_state(self.h, i, _dp(a)), self.h)
        return a[:, BF:-BF, BF:-BF].copy()

    def run(self, nsteps: int, want_dts: bool = False):
        t = np.zeros(self.n)
        hist = np.zeros((max(1, nsteps), self.n)) if want_dts else None
        self._check(self.lib.dense_batch_run(self.h, nsteps, _dp(hist) if want_dts else None, _dp(t)), self.h)
        self.t = t
        return hist[:nsteps] if want_dts else None

    def set_schedule(self, tile: int, order: int = 0, grid_pct: int = 100):
        self._check(self.lib.dense_batch_set_schedule(self.h, tile, order, grid_pct), self.h)

    def autotune(self, trial_steps: int = 4):
        t, o, g = C.c_int(), C.c_int(), C.c_int()
        ms = C.c_double()
        self._check(self.lib.dense_batch_autotune(self.h, trial_steps, C.byref(t), C.byref(o), C.byref(g),
                                                  C.byref(ms)), self.h)
        return t.value, o.value, g.value, ms.value

    def last_run_gpu_ms(self) -> float:
        ms = C.c_double()
        self._check(self.lib.dense_batch_last_run_gpu_ms(self.h, C.byref(ms)), self.h)
        return ms.value


# ---------------------------------------------------------------- ensemble

def cons_from_prim(rho, u, v, p, gamma):
    return np.stack([rho, rho * u, rho * v, p / (gamma - 1.0) + 0.5 * rho * (u * u + v * v)])


def vortex_state(n: int, eps: float, xc: float, yc: float, gamma: float = 1.4, L: float = 10.0,
                 uinf: float = 1.0, vinf: float = 1.0):
    """Isentropic vortex of strength eps at (xc, yc) in a free stream
    rho = p = 1, (u, v) = (uinf, vinf) on the periodic square [0, L]^2
    (minimum-image distance)."""
    h = L / n
    x = (np.arange(n) + 0.5) * h
    X, Y = np.meshgrid(x, x)
    dxp = (X - xc + 0.5 * L) % L - 0.5 * L
    dyp = (Y - yc + 0.5 * L) % L - 0.5 * L
    r2 = dxp * dxp + dyp * dyp
    f = eps / (2.0 * math.pi) * np.exp(0.5 * (1.0 - r2))
    u = uinf - f * dyp
    v = vinf + f * dxp
    T = 1.0 - (gamma - 1.0) * eps * eps * np.exp(1.0 - r2) / (8.0 * gamma * math.pi ** 2)
    rho = T ** (1.0 / (gamma - 1.0))
    p = rho ** gamma
    return cons_from_prim(rho, u, v, p, gamma), h


def shock_post_state(mach: float, gamma: float = 1.4):
    """Rankine-Hugoniot state behind a normal shock of Mach `mach` running into
    rho = 1, p = 1 with the shock at rest (upstream u = mach*sqrt(gamma))."""
    u1 = mach * math.sqrt(gamma)
    rho2 = (gamma + 1.0) * mach ** 2 / (2.0 + (gamma - 1.0) * mach ** 2)
    p2 = 1.0 + 2.0 * gamma / (gamma + 1.0) * (mach ** 2 - 1.0)
    u2 = (2.0 + (gamma - 1.0) * mach ** 2) / ((gamma + 1.0) * mach ** 2) * u1
    return rho2, u2, p2


def shock_vortex_state(nx: int, ny: int, mach: float, xc: float, yc: float, gamma: float = 1.4,
                       eps: float = 0.3, rc: float = 0.05, alpha: float = 0.204, xs: float = 0.5):
    """Stationary normal shock at x = xs on [0,2]x[0,1] with a compressible
    vortex (eps, rc, alpha) advected by the upstream flow toward it."""
    h = 2.0 / nx
    x = (np.arange(nx) + 0.5) * h
    y = (np.arange(ny) + 0.5) * (1.0 / ny)
    X, Y = np.meshgrid(x, y)
    r = np.sqrt((X - xc) ** 2 + (Y - yc) ** 2)
    r = np.maximum(r, 1e-12)
    tau = r / rc
    rho_v = (1.0 - (gamma - 1.0) * eps ** 2 * np.exp(2 * alpha * (1 - tau ** 2)) / (4 * alpha * gamma)) ** (
        1.0 / (gamma - 1.0))
    sw = eps * tau * np.exp(alpha * (1.0 - tau ** 2))
    u_v = mach * math.sqrt(gamma) + sw * (Y - yc) / r
    v_v = -sw * (X - xc) / r
    p_v = rho_v ** gamma
    rho2, u2, p2 = shock_post_state(mach, gamma)
    left = X < xs
    rho = np.where(left, rho_v, rho2)
    u = np.where(left, u_v, u2)
    v = np.where(left, v_v, 0.0)
    p = np.where(left, p_v, p2)
    return cons_from_prim(rho, u, v, p, gamma), h, 1.0 / ny


def ic_desc(inst) -> IcDesc:
    """Device initial-state descriptor of an ensemble instance (the same
    flow ``vortex_state`` / ``shock_vortex_state`` build on the host)."""
    d = IcDesc()
    P = inst.params
    if inst.kind == "vortex":
        d.kind, d.eps, d.xc, d.yc = IC_VORTEX, P["eps"], P["xc"], P["yc"]
        d.uinf, d.vinf, d.length = P.get("uinf", 1.0), P.get("vinf", 1.0), P.get("L", 10.0)
    elif inst.kind == "shock_vortex":
        d.kind, d.mach, d.xc, d.yc = IC_SHOCK_VORTEX, P["mach"], P["xc"], P["yc"]
        d.eps, d.rc, d.alpha, d.xs = P.get("eps", 0.3), P.get("rc", 0.05), P.get("alpha", 0.204), P.get("xs", 0.5)
    elif inst.kind == "shock_vortex_ref":  # the reference's benchmark 9, no free parameters
        d.kind = IC_SHOCK_VORTEX_REF
    else:
        raise ValueError(f"unknown instance kind {inst.kind!r}")
    return d


@dataclass
class Instance:
    kind: str          # "vortex", "shock_vortex" or "shock_vortex_ref"
    ni: int
    nj: int
    params: dict


def make_ensemble(n: int = 256, seed: int = 2026, nmin: int = 64, nmax: int = 512, splitting: int = LLF,
                  recon: int = WENOZ, cfl: float = 0.5, tile: int = 0, host_states: bool = True):
    """Seeded ensemble: half isentropic vortices (strength U[1,5], centre
    U(domain), periodic square), half shock/vortex interactions (shock Mach
    U[1.05,1.5], vortex centre U in the upstream region, extrapolation in x,
    reflective walls in y).  Grid sides are log-uniform in [nmin, nmax]
    (shock/vortex grids are 2:1).  Returns (cfgs, states, instances);
    with host_states=False the states are None (generate them on the device
    with ``Batch.init_ic(instances)``)."""
    rng = np.random.default_rng(seed)
    cfgs, states, insts = [], [], []
    for k in range(n):
        side = int(round(nmin * (nmax / nmin) ** rng.random()))
        side = max(8, side)
        if k % 2 == 0:
            eps = rng.uniform(1.0, 5.0)
            xc, yc = rng.uniform(0.0, 10.0, size=2)
            h = 10.0 / side
            cons = vortex_state(side, eps, xc, yc)[0] if host_states else None
            cfg = make_cfg(side, side, benchmark=0, splitting=splitting, recon=recon, cfl=cfl, gamma=1.4,
                           dx=h, dy=h, bc=[BC_PERIODIC] * 4, tile=tile)
            insts.append(Instance("vortex", side, side, dict(eps=eps, xc=xc, yc=yc)))
        else:
            mach = rng.uniform(1.05, 1.5)
            xc = rng.uniform(0.15, 0.35)
            yc = rng.uniform(0.25, 0.75)
            nx, ny = side, max(8, side // 2)
            hx, hy = 2.0 / nx, 1.0 / ny
            cons = shock_vortex_state(nx, ny, mach, xc, yc)[0] if host_states else None
            cfg = make_cfg(nx, ny, benchmark=0, splitting=splitting, recon=recon, cfl=cfl, gamma=1.4,
                           dx=hx, dy=hy, bc=[BC_EXTRAP, BC_EXTRAP, BC_REFLECT, BC_REFLECT], tile=tile)
            insts.append(Instance("shock_vortex", nx, ny, dict(mach=mach, xc=xc, yc=yc)))
        cfgs.append(cfg)
        states.append(cons)
    return cfgs, states, insts
