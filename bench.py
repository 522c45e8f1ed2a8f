#!/usr/bin/env python3
"""Benchmark: FP64 cell-updates/s of the explicit update on one MI355X.

Workload (one "step" = one full 3-stage SSP-RK3 time step of the hot path):
double Mach reflection, 1920 x 480 cells on [0,4]x[0,1], Mach 10 shock at
60 deg, characteristic WENO-Z with global Lax-Friedrichs splitting
(switch 104), RK3, CFL 0.6, gamma 1.4, FP64.  The state starts from the
problem's initial condition and stays resident in HBM; the timed region is
K steps replayed from captured HIP graphs with the clock on the device,
bracketed by device synchronisation.

  value   = 1920*480*K / wall time of the K steps        (cell-updates/s)
  roofline: dominant kernel = the fused RK-stage kernel; achieved =
            algorithmic FP64 flops per launch / average launch duration
            (HIP events on the library's stream over the timed region);
            peak = MI355X FP64 vector 78.6 TFLOP/s.
  cpu_baseline: the reference solver compiled from its sources
            (oracle/_ref/2dense_ref_racefix: the OpenMP build with the
            eta-sweep race fixed), run on the host cores on a bounded
            sample of the same workload, timed by its own CpuTime.dat.

  batched: 256 seeded vortex and shock/vortex blocks (64^2..512^2, LLF +
            WENO-Z, 200 steps each): sequential plain GPU path vs one
            persistent launch per RK stage for the whole ensemble.
  adjoint: gradient of a density mismatch after 200 steps (256^2 periodic
            vortex, 64 Fourier modes, LLF + WENO-Z, checkpoint every 20 steps)
            vs a plain forward solve of the same problem.
  other_workloads: GPU cell-updates/s of the other configured workloads
            (isentropic vortex 200^2, double Mach 960x240, Rayleigh-Taylor
            256x1024 and 1024x4096); reported beside the metric, not in it.

    python bench.py [--gpus 1] [--steps 640] [--warmup 32] [--no-cpu] [--no-batched] [--no-adjoint]
                    [--no-workloads]

Steps are replayed in graph chunks of 16; multiples of 16 keep every
recorded stage launch a working one.
"""
from __future__ import annotations

import argparse
import importlib
import json
import os
import shutil
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

NI, NJ = 1920, 480
SPLITTING, RECON, CFL = 104, 255, 0.6
# Algorithmic FP64 ops (+,-,*,/,sqrt; no FMA fusion) per cell-update of this
# workload, counted by the oracle's op-counting build: python oracle/count_ops.py 5 1920 480 104 255
FLOPS_PER_CELL_UPDATE = 6063.015625
# State traffic of a fused ping-pong RK3 step: 4 FP64 variables,
# read U^n / write U1; read U^n,U1 / write U2; read U^n,U2 / write U^n+1.
BYTES_PER_CELL_UPDATE = 256.0
# The same count for every switch pair on double Mach 1920x480 (first step from
# the initial condition; AdaWENO's depends on how many faces take the
# characteristic branch): python oracle/count_ops.py 5 1920 480 <splitting> <recon>
SCHEME_FLOPS = {(101, 251): 794.21875, (101, 253): 3677.96875, (101, 255): 4014.4,
                (102, 251): 788.2, (102, 253): 3671.90625, (102, 255): 4008.34375,
                (103, 251): 1412.59375, (103, 253): 4296.34375, (103, 255): 4632.8,
                (104, 255): FLOPS_PER_CELL_UPDATE, (105, 255): 2026.7}
PEAK_FP64_VALU = 78.6e12
PEAK_HBM = 8.0e12


def dmr_initial_state(ni: int, nj: int, gamma: float = 1.4) -> np.ndarray:
    """Interior conserved state of the double Mach reflection
    (init_doublemach_general.f90:36-59): post-shock (8, 8.25 cos30, -8.25 sin30, 116.5)
    left of x = 1/6 + y/sqrt(3), pre-shock (1.4, 0, 0, 1) right of it."""
    dx, dy = 4.0 / ni, 1.0 / nj
    x = (np.arange(1, ni + 1, dtype=np.float64) - 0.5) * dx
    y = (np.arange(1, nj + 1, dtype=np.float64) - 0.5) * dy
    X, Y = np.meshgrid(x, y)
    post = X < (1.0 / 6.0 + Y * np.sqrt(1.0 / 3.0))
    rho = np.where(post, 8.0, 1.4)
    u = np.where(post, 8.25 * np.sqrt(3.0) * 0.5, 0.0)
    v = np.where(post, -8.25 * 0.5, 0.0)
    p = np.where(post, 116.5, 1.0)
    cons = np.empty((4, nj, ni))
    cons[0] = rho
    cons[1] = rho * u
    cons[2] = rho * v
    cons[3] = p / (gamma - 1.0) + 0.5 * rho * (u * u + v * v)
    return cons


def rt_initial_state(ni: int, nj: int, gamma: float = 5.0 / 3.0) -> np.ndarray:
    """Interior conserved state of the Rayleigh-Taylor problem
    (init_RT_general.f90:16-48): node-centred grid on [0,0.25]x[0,1], heavy
    fluid rho = 2 below y = 0.5 with p = 2y + 1, light fluid rho = 1 above with
    p = y + 1.5, v = -0.025 c cos(8 pi x) (the light side's literal is single
    precision, :47)."""
    dx, dy = 0.25 / (ni - 1.0), 1.0 / (nj - 1.0)
    x = np.arange(ni) * dx
    y = np.arange(nj) * dy
    X, Y = np.meshgrid(x, y)
    heavy = Y <= 0.5
    rho = np.where(heavy, 2.0, 1.0)
    p = np.where(heavy, 2.0 * Y + 1.0, Y + 1.5)
    c = np.sqrt(gamma * p / rho)
    amp = np.where(heavy, -0.025, float(np.float32(-0.025)))
    v = amp * c * np.cos(8.0 * np.pi * X)
    return np.stack([rho, np.zeros_like(rho), rho * v, p / (gamma - 1.0) + 0.5 * rho * v * v])


TILE_NAMES = {1: "32x7", 2: "8x7", 3: "32x15", 4: "16x15", 5: "16x15/512"}


def workloads_leg(reps: int = 3):
    """GPU throughput of the other configured workloads (parity cases, not the
    metric): tile shape from dense_autotune, then a first run that captures the
    graphs and `reps` timed runs of the same steps from the same initial state;
    best wall time, cell-updates/s."""
    D = importlib.import_module("2dense_amd")
    B = importlib.import_module("2dense_amd.batch")
    vort, h = B.vortex_state(200, 5.0, 5.0, 5.0)
    cases = [
        ("isentropic vortex 200x200, LLF + WENO-JS (101/253), CFL 0.5, periodic",
         D.make_cfg(200, 200, benchmark=0, splitting=101, recon=253, cfl=0.5, gamma=1.4, dx=h, dy=h,
                    bc=[D.BC_PERIODIC] * 4), vort, 100),
        ("double Mach reflection 960x240, characteristic WENO-Z + GLF (104/255), CFL 0.6",
         D.make_cfg(960, 240, benchmark=5, splitting=104, recon=255, cfl=0.6), dmr_initial_state(960, 240), 320),
        ("Rayleigh-Taylor 256x1024, AdaWENO (105/255), gravity, CFL 0.5",
         D.make_cfg(256, 1024, benchmark=3, splitting=105, recon=255, cfl=0.5), rt_initial_state(256, 1024), 320),
        ("Rayleigh-Taylor 1024x4096, AdaWENO (105/255), gravity, CFL 0.5",
         D.make_cfg(1024, 4096, benchmark=3, splitting=105, recon=255, cfl=0.5), rt_initial_state(1024, 4096), 64),
    ]
    out = []
    for name, cfg, cons, steps in cases:
        best = None
        with D.Solver(cfg) as s:
            s.set_state(cons)
            tile, _ = s.autotune(16)  # the same bits for every shape
            for r in range(reps + 1):
                s.set_state(cons)
                s.current_time, s.current_step = 0.0, 0
                s.sync()
                t0 = time.perf_counter()
                n = s.run(steps)
                s.sync()
                wall = time.perf_counter() - t0
                if n != steps:
                    raise RuntimeError(f"{name}: ran {n} of {steps} steps")
                if r > 0 and (best is None or wall < best[0]):
                    best = (wall, s.last_run_gpu_ms()[0])
        cells = cfg.ni * cfg.nj
        out.append(dict(workload=name, ni=cfg.ni, nj=cfg.nj, steps=steps, tile=TILE_NAMES[tile],
                        cell_updates_per_s=cells * steps / best[0], gpu_cell_updates_per_s=cells * steps / (best[1] * 1e-3),
                        ms_per_step=best[0] / steps * 1e3))
    return out


def scheme_table(steps: int = 160):
    """Every switch pair on double Mach 1920x480 from the initial condition with
    the automatic tile shape: per pair, us per stage launch (HIP events on the
    library's stream), cell-updates/s (wall clock over graph-replayed steps, best
    of two timed runs after one that captures the graphs), and the algorithmic
    FP64 and HBM rates of the stage kernel against MI355X's peaks."""
    D = importlib.import_module("2dense_amd")
    cons = dmr_initial_state(NI, NJ)
    rows = []
    for (sp, rc), flops in SCHEME_FLOPS.items():
        best = None
        with D.Solver(D.make_cfg(NI, NJ, benchmark=5, splitting=sp, recon=rc, cfl=CFL)) as s:
            for r in range(3):
                s.set_state(cons)
                s.current_time, s.current_step = 0.0, 0
                s.sync()
                t0 = time.perf_counter()
                n = s.run(steps)
                s.sync()
                wall = time.perf_counter() - t0
                if n != steps:
                    raise RuntimeError(f"{sp}/{rc}: ran {n} of {steps} steps")
                ms, launches = s.last_run_gpu_ms()
                if r > 0 and (best is None or wall < best[0]):
                    best = (wall, ms / launches * 1e-3)
        wall, t_launch = best
        cells_launch = NI * NJ
        tflops = flops / 3.0 * cells_launch / t_launch
        gbs = BYTES_PER_CELL_UPDATE / 3.0 * cells_launch / t_launch
        rows.append(dict(splitting=sp, recon=rc, us_per_launch=t_launch * 1e6, cell_updates_per_s=cells_launch * steps / wall,
                         flops_per_cell_update=flops, algorithmic_tflops=tflops / 1e12, fp64_frac=tflops / PEAK_FP64_VALU,
                         algorithmic_hbm_gbs=gbs / 1e9, hbm_frac=gbs / PEAK_HBM))
    return rows


def gpu_leg(steps: int, warmup: int):
    D = importlib.import_module("2dense_amd")
    cfg = D.make_cfg(NI, NJ, benchmark=5, splitting=SPLITTING, recon=RECON, cfl=CFL, tile=4)  # 16x15
    s = D.Solver(cfg)
    s.set_state(dmr_initial_state(NI, NJ))
    if warmup:
        s.run(warmup)
    if steps % 16:
        s.run(steps % 16)  # captures the remainder graph of the timed run outside the timed region
    s.sync()
    t0 = time.perf_counter()
    taken = s.run(steps)
    s.sync()
    wall = time.perf_counter() - t0
    gpu_ms, launches = s.last_run_gpu_ms()
    state = s.interior()
    if not np.all(np.isfinite(state)):
        raise RuntimeError("non-finite state after the timed run")
    s.close()
    return taken, wall, gpu_ms, launches


def batched_leg(n: int = 256, steps: int = 200, seed: int = 2026):
    """Ensemble of n independent blocks (config: isentropic vortices with
    strength U[1,5] and shock/vortex interactions with Mach U[1.05,1.5], sides
    log-uniform 64..512, LLF + WENO-Z, RK3, CFL 0.5, `steps` steps each):
    sequential plain GPU path (one Solver per instance, graph-replayed steps)
    vs the batched path (one persistent launch per RK stage for all instances,
    schedule from the autotuner).  Only the stepping is timed in both: each
    path first runs 16 + steps % 16 steps untimed, which captures the graphs
    the timed run replays (a 16-step graph and one of the remainder), then
    starts again from the initial states."""
    D = importlib.import_module("2dense_amd")
    B = importlib.import_module("2dense_amd.batch")
    cfgs, states, insts = B.make_ensemble(n=n, seed=seed)
    cells = sum(c.ni * c.nj for c in cfgs)
    warm = 16 + steps % 16
    seq_wall = 0.0
    for cfg, cons in zip(cfgs, states):
        s = D.Solver(cfg)
        s.set_state(cons)
        s.run(warm)
        s.set_state(cons)
        s.current_time, s.current_step = 0.0, 0
        s.sync()
        t0 = time.perf_counter()
        s.run(steps)
        s.sync()
        seq_wall += time.perf_counter() - t0
        s.close()
    b = B.Batch(cfgs)
    for i, cons in enumerate(states):
        b.set_state(i, cons)
    tile, order, grid_pct, _ = b.autotune(trial_steps=4)
    b.run(warm)
    for i, cons in enumerate(states):
        b.set_state(i, cons)
    t0 = time.perf_counter()
    b.run(steps)
    bat_wall = time.perf_counter() - t0
    b.close()
    return dict(instances=n, steps=steps, cells=cells, seed=seed,
                sequential_cell_updates_per_s=cells * steps / seq_wall,
                batched_cell_updates_per_s=cells * steps / bat_wall,
                speedup=seq_wall / bat_wall, sequential_s=seq_wall, batched_s=bat_wall,
                schedule={"tile": tile, "order": order, "grid_pct": grid_pct})


def adjoint_leg(n: int = 256, steps: int = 200, k: int = 20, n_modes: int = 64, fd_modes: int = 8,
                reps: int = 3):
    """Gradient of J = sum (rho_N - rho*)^2 dx dy with respect to 64 Fourier
    amplitudes of the initial density (periodic 256x256 vortex, LLF + WENO-Z,
    RK3, CFL 0.4, 200 steps, checkpoint every 20 steps) through the
    torch.autograd.Function, against a plain forward solve of the same problem
    (graph-replayed Solver.run).  Reports the cost ratio and a central
    finite-difference spot check (h = 1e-6) on the first fd_modes amplitudes."""
    import torch
    D = importlib.import_module("2dense_amd")
    A = importlib.import_module("2dense_amd.adjoint")
    ic = A.FourierIC(n, n_modes=n_modes, device="cuda")
    theta = torch.tensor(A.draw_amplitudes(n_modes), dtype=torch.float64, device="cuda")
    target = ic(torch.tensor(A.draw_amplitudes(n_modes, seed=12), dtype=torch.float64, device="cuda"))[0]
    # plain forward solve
    cfg = D.make_cfg(n, n, benchmark=0, splitting=D.LLF, recon=D.WENOZ, cfl=0.4, dx=ic.h, dy=ic.h,
                     bc=[D.BC_PERIODIC] * 4)
    U0 = ic(theta).cpu().numpy()
    plain_wall, plain_gpu = [], []
    with D.Solver(cfg) as s:
        for _ in range(reps + 1):
            s.set_state(U0)
            s.current_time, s.current_step = 0.0, 0
            s.sync()
            t0 = time.perf_counter()
            s.run(steps)
            s.sync()
            plain_wall.append(time.perf_counter() - t0)
            plain_gpu.append(s.last_run_gpu_ms()[0])
    # gradient: forward with checkpoints + backward, through autograd
    R = A.Rollout(n, n, steps, dx=ic.h, dy=ic.h, cfl=0.4, checkpoint_every=k)
    grad_wall, fwd_gpu, bwd_gpu = [], [], []
    g = None
    for _ in range(reps + 1):
        th = theta.clone().requires_grad_(True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        J = A.density_mismatch(A.rollout(ic(th), R), target, ic.h, ic.h)
        J.backward()
        torch.cuda.synchronize()
        grad_wall.append(time.perf_counter() - t0)
        f, b = R.last_ms()
        fwd_gpu.append(f)
        bwd_gpu.append(b)
        g = th.grad.cpu().numpy()
    dts = R.dts.copy()

    def Jat(t):
        uN, _ = R.forward(ic(t), dt_schedule=dts)
        return float(A.density_mismatch(uN, target, ic.h, ic.h))

    fd = np.zeros(fd_modes)
    h = 1e-6
    for m in range(fd_modes):
        e = torch.zeros_like(theta)
        e[m] = h
        fd[m] = (Jat(theta + e) - Jat(theta - e)) / (2 * h)
    R.close()
    med = lambda x: float(np.median(x[1:]))
    return dict(workload=f"periodic vortex {n}x{n}, {n_modes} Fourier density modes, LLF + WENO-Z (101/255), "
                         f"SSP-RK3, CFL 0.4, {steps} steps, checkpoint every {k}",
                plain_forward_ms=med(plain_wall) * 1e3, plain_forward_gpu_ms=med(plain_gpu),
                gradient_ms=med(grad_wall) * 1e3, adjoint_forward_gpu_ms=med(fwd_gpu),
                adjoint_backward_gpu_ms=med(bwd_gpu),
                cost_ratio=med(grad_wall) / med(plain_wall),
                cost_ratio_gpu=(med(fwd_gpu) + med(bwd_gpu)) / med(plain_gpu),
                fd_check=dict(modes=fd_modes, h=h,
                              l2_rel=float(np.linalg.norm(g[:fd_modes] - fd) / np.linalg.norm(fd)),
                              worst_rel=float(np.max(np.abs(g[:fd_modes] - fd) / np.abs(fd)))))


def probe_leg():
    """FP64 FMA rate and HBM streaming rates measured on this device
    (dense_probe_fp64 / dense_probe_hbm over 2 GiB, beyond the last-level cache)."""
    D = importlib.import_module("2dense_amd")
    return D.probe_peaks(2 << 30)


def driver_leg(total_time: float = 0.2):
    """The Fortran drop-in driver (2dense_amd/bin/dense_gpu, the reference's
    input.in and loop, general_main.f90:23-38) on the headline workload as
    configured, to t = 0.2, timed by its own CpuTime.dat as the reference
    times its loop (graph captures at the first steps included).  It replays
    the steps between saves from captured graphs (dense_run_saving)."""
    sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))
    drv = os.path.join(ROOT, "2dense_amd", "bin", "dense_gpu")
    if not os.path.exists(drv):
        return dict(error="Fortran driver not built")
    wd = tempfile.mkdtemp(prefix="dense_drv_")
    try:
        with open(os.path.join(wd, "input.in"), "w") as f:
            f.write(_input_in(benchmark=5, ni=NI, nj=NJ, splitting=SPLITTING, recon=RECON, cfl=CFL,
                              total_step=10 ** 8, save_step=10 ** 8, threads=1, total_time=total_time))
        env = dict(os.environ, DENSE_AUTOTUNE="0")
        out = subprocess.run([drv], cwd=wd, env=env, capture_output=True, text=True, timeout=600)
        if out.returncode != 0:
            return dict(error=out.stdout[-500:] + out.stderr[-500:])
        with open(os.path.join(wd, "CpuTime.dat")) as f:
            fields = f.readline().split(":")[1].split()
        t, steps = float(fields[0].replace("D", "E")), int(fields[1])
        return dict(workload=f"double Mach {NI}x{NJ}, switch {SPLITTING}, CFL {CFL}, to t = {total_time} through the "
                             "Fortran driver", steps=steps, loop_s=t, cell_updates_per_s=NI * NJ * steps / t)
    finally:
        shutil.rmtree(wd, ignore_errors=True)


def _input_in(*, benchmark, ni, nj, splitting, recon, cfl, total_step, save_step, threads, gamma=1.4,
              total_time=1.0e30):
    """The reference's six NAMELIST groups (flow_mod_general.f90:313-332) with
    output off: no dumps, no Tecplot saves before the end of the run."""
    return f"""&OMP_PARS
 OMP_THREAD_NUM={threads},
 /
&MODEL_PARS
 NEQ=4, IF_SOURCE=F, IF_VISCOUS=F, IF_BODYFORCE=F, MACH_INF=1.0, REYNOLDS=500.0,
 BENCHMARK_INDEX={benchmark},
 /
&GRID_PARS
 IF_CURVE_GRID=F, IF_READ_FROM_FILE=F, SINGLE_NI={ni}, SINGLE_NJ={nj},
 /
&TIME_STEP_PARS
 STAGES=3, TOTAL_TIME={total_time!r}, DT_SAVE=1.0e30, DT=0.01, CFL={cfl!r},
 TOTAL_STEP={total_step}, SAVE_STEP={save_step},
 IF_CONST_DT=F, IF_SAVE_BINARY=F, IF_SAVE_TEC=F,
 /
&SCHEME_PARS
 SCHEME_SPLITTING_INDEX={splitting}, SCHEME_INV_INDEX={recon}, SCHEME_VIS_INDEX=301,
 IF_SAVE_THETA=F, DELTA=0.5,
 /
&GAS_PARS
 GAMMA={gamma!r},
 /
"""


# AdaWENO's algorithmic flops per cell-update with every (face, sign) pair on
# the component-wise branch (delta = 0) and on the characteristic branch
# (delta = 1), from the oracle's op count: python oracle/count_ops.py <bm> <ni> <nj> 105 255 <delta>.
# The reference's cost is linear in the fraction of characteristic pairs in between.
ADAWENO_FLOPS = {"dmr": (2013.859375, 6808.09375),            # 5 1920 480 105 255 {0,1}
                 "rt": (2035.80908203125, 6826.7314453125)}    # 3 1024 4096 105 255 {0,1}


def _branch_stats(th, delta, ni, nj, tx=16, ty=15, nt=256):
    """Fraction of (face, sign) pairs on the characteristic branch (theta <=
    delta) and the fraction of wavefronts, per direction and sign, whose 64
    faces take both branches, for the automatic 16x15 tile with 256 threads
    (x faces: thread t -> face (t % 17, t // 17); y faces: (t % 16, t // 16))."""
    xp, xn, ep, en = th
    nchar = sum(int(np.count_nonzero(a <= delta)) for a in th)
    npairs = sum(a.size for a in th)
    tiles_x, tiles_y = -(-ni // tx), -(-nj // ty)
    t = np.arange(nt)
    X0 = (np.arange(tiles_x) * tx)[None, :, None]
    Y0 = (np.arange(tiles_y) * ty)[:, None, None]
    mixed, waves = 0, 0
    for d, arrs in ((0, (xp, xn)), (1, (ep, en))):
        if d == 0:
            i, j = X0 + t % (tx + 1), Y0 + t // (tx + 1)
            valid = (t < (tx + 1) * ty) & (i <= ni) & (j < nj)
        else:
            i, j = X0 + t % tx, Y0 + t // tx
            valid = (t < tx * (ty + 1)) & (i < ni) & (j <= nj)
        ic, jc = np.minimum(i, ni if d == 0 else ni - 1), np.minimum(j, nj - 1 if d == 0 else nj)
        for a in arrs:
            v = a[jc, ic] if d == 0 else a[ic, jc]
            ch = (v <= delta) & valid
            cw = (v > delta) & valid
            shp = ch.shape[:2] + (nt // 64, 64)
            anych, anycw = ch.reshape(shp).any(-1), cw.reshape(shp).any(-1)
            mixed += int(np.count_nonzero(anych & anycw))
            waves += int(np.count_nonzero(anych | anycw))
    return nchar / npairs, mixed / waves


def adaweno_leg(delta: float = 0.5, reps: int = 2):
    """AdaWENO (105) on developed flows: Rayleigh-Taylor 1024x4096 run to
    t = 1.0 (the mixing layer) and double Mach 1920x480 run to t = 0.1.  At
    each state: the fraction of (face, sign) pairs on the characteristic
    branch (theta <= delta, dense_get_theta), the fraction of wavefronts that
    mix both branches, us per stage launch at delta = 0 (all component-wise),
    the run's delta and 1 (all characteristic) from the same state, and the
    divergence cost: the measured time over the time a perfectly sorted mix
    of the two branches would take, T0 + f (T1 - T0)."""
    D = importlib.import_module("2dense_amd")
    cases = [("rt", "Rayleigh-Taylor 1024x4096 at t = 1.0", D.make_cfg(1024, 4096, benchmark=3, splitting=105,
                                                                       recon=255, cfl=0.5, delta=delta),
              rt_initial_state(1024, 4096), 1.0),
             ("dmr", "double Mach 1920x480 at t = 0.1", D.make_cfg(NI, NJ, benchmark=5, splitting=105, recon=255,
                                                                  cfl=CFL, delta=delta),
              dmr_initial_state(NI, NJ), 0.1)]
    out = []
    for key, name, cfg, ic, t_dev in cases:
        with D.Solver(cfg) as s:
            s.set_state(ic)
            n_dev = s.run(10 ** 7, total_time=t_dev)
            th = s.theta()
            frac, mixed = _branch_stats(th, delta, cfg.ni, cfg.nj)
            state, t0, st0 = s.interior(), s.current_time, s.current_step
            us = {}
            for d in (0.0, delta, 1.0):
                s.update_params(delta=d)
                best = None
                for _ in range(reps + 1):
                    s.set_state(state)
                    s.current_time, s.current_step = t0, st0
                    s.run(16)
                    ms, n = s.last_run_gpu_ms()
                    best = ms / n * 1e3 if best is None else min(best, ms / n * 1e3)
                us[d] = best
        ideal = us[0.0] + frac * (us[1.0] - us[0.0])
        row = dict(workload=name, steps_to_state=n_dev, char_pair_fraction=frac, mixed_wave_fraction=mixed,
                   us_per_launch=us[delta], us_all_componentwise=us[0.0], us_all_characteristic=us[1.0],
                   divergence_cost=us[delta] / ideal - 1.0,
                   cell_updates_per_s=cfg.ni * cfg.nj / (3 * us[delta] * 1e-6))
        f0, f1 = ADAWENO_FLOPS[key]
        flops = f0 + frac * (f1 - f0)
        row.update(flops_per_cell_update=flops,
                   algorithmic_tflops=flops / 3.0 * cfg.ni * cfg.nj / (us[delta] * 1e-6) / 1e12)
        out.append(row)
    return out


def _host_info() -> dict:
    info = dict(nproc=os.cpu_count())
    try:
        info["affinity"] = len(os.sched_getaffinity(0))
    except AttributeError:
        pass
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    info["model"] = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return info


def _threads_available() -> int:
    """Host threads for the CPU baseline: the cores this process may use, at
    most the 16 host cores a one-GPU lease is allotted (the GPU box shows more
    cores than the lease owns)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(n, 16))


def cpu_leg(threads: int, steps: int):
    """Time the reference binary on the host cores (CpuTime.dat timer)."""
    ref = os.path.join(ROOT, "oracle", "_ref", "2dense_ref_racefix")
    if os.path.exists(ref):
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pyoracle as po
        wd = tempfile.mkdtemp(prefix="dense_cpu_")
        try:
            po.write_input_in(os.path.join(wd, "input.in"), benchmark=5, ni=NI, nj=NJ, splitting=SPLITTING,
                              recon=RECON, cfl=CFL, total_step=steps, save_step=10 ** 8, threads=threads,
                              if_save_binary=False, if_save_tec=False)
            env = dict(os.environ, OMP_PROC_BIND="close", OMP_PLACES="cores")
            subprocess.run([ref], cwd=wd, env=env, capture_output=True, text=True, check=True, timeout=600)
            with open(os.path.join(wd, "CpuTime.dat")) as f:
                line = f.readline()
            cpu_t = float(line.split(":")[1].split()[0].replace("D", "E"))
            return dict(value=NI * NJ * steps / cpu_t, unit="cell-updates/s", cores=threads, kind="reference",
                        sample=f"reference solver (flang -O2 -fopenmp, eta-sweep race fixed), DMR {NI}x{NJ} "
                               f"switch {SPLITTING}, {steps} steps, {threads} OpenMP threads, "
                               f"timed by its own CpuTime.dat ({cpu_t:.2f} s)", host=_host_info())
        finally:
            shutil.rmtree(wd, ignore_errors=True)
    # fallback: the C restatement (bitwise equal to the reference)
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as po
    p = po.problem_params(5, NI, NJ, splitting=SPLITTING, recon=RECON, cfl=CFL, nthreads=threads)
    o = po.Oracle(p)
    o.init_state(5)
    t0 = time.perf_counter()
    o.step(steps)
    dt = time.perf_counter() - t0
    return dict(value=NI * NJ * steps / dt, unit="cell-updates/s", cores=threads, kind="port",
                sample=f"C restatement, DMR {NI}x{NJ} switch {SPLITTING}, {steps} steps, {threads} threads")


def hbm_probe(nbytes: int):
    """The streaming read and copy kernels of dense_probe_hbm over nbytes, to
    calibrate the HBM counters: per dispatch, read_probe_kernel fetches nbytes
    and copy_probe_kernel fetches and writes nbytes each."""
    D = importlib.import_module("2dense_amd")
    print(json.dumps(D.probe_peaks(nbytes)))


def read_traffic():
    """HBM bytes per stage launch from the committed rocprofv3 --pmc summary, if any."""
    p = os.path.join(ROOT, "profiles", "stage_kernel_traffic.json")
    if os.path.exists(p):
        with open(p) as f:
            return json.load(f).get("hbm_bytes_per_launch")
    return None


def _leg(name, fn, *a):
    """Run one leg and report it on stderr (stdout carries only the JSON line)."""
    t0 = time.time()
    r = fn(*a)
    print(f"bench.py: {name} done in {time.time() - t0:.1f} s", file=sys.stderr, flush=True)
    return r


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=640)
    ap.add_argument("--warmup", type=int, default=32)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=64)  # ~14 s of the reference on 16 host cores
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-workloads", action="store_true")
    ap.add_argument("--scheme-table", action="store_true",
                    help="only time every switch pair on double Mach 1920x480 (one JSON line per pair)")
    ap.add_argument("--no-driver", action="store_true")
    ap.add_argument("--no-adaweno", action="store_true")
    ap.add_argument("--hbm-probe", type=int, default=0, metavar="BYTES",
                    help="only run the streaming read/copy probe over BYTES (HBM counter calibration)")
    args = ap.parse_args()
    if args.hbm_probe:
        hbm_probe(args.hbm_probe)
        return
    if args.scheme_table:
        for row in scheme_table():
            print(json.dumps(row), flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(multi-GPU sharding is design-only)", file=sys.stderr)
        sys.exit(2)

    taken, wall, gpu_ms, launches = _leg("double Mach 1920x480", gpu_leg, args.steps, args.warmup)
    if taken != args.steps:
        raise RuntimeError(f"ran {taken} of {args.steps} steps")
    cu = NI * NJ * taken
    value = cu / wall
    ms_launch = gpu_ms / max(launches, 1)
    flops_launch = FLOPS_PER_CELL_UPDATE * NI * NJ / 3.0
    achieved = flops_launch / (ms_launch * 1e-3)
    peaks = _leg("roofline probes", probe_leg)
    # the roofline's peak: the lower of MI355X's published FP64 rate and the
    # FMA-loop rate measured on this device
    peak = min(PEAK_FP64_VALU, peaks["fp64_tflops"] * 1e12)
    hbm_peak = min(PEAK_HBM, peaks["hbm_read_gbs"] * 1e9)
    hbm_alg = BYTES_PER_CELL_UPDATE / 3.0 * NI * NJ / (ms_launch * 1e-3)
    roofline = dict(bound="fp64-valu", achieved=achieved / 1e12, peak=peak / 1e12, unit="TFLOP/s",
                    frac=achieved / peak, traffic=read_traffic(),
                    peak_published=PEAK_FP64_VALU / 1e12, peak_measured=peaks["fp64_tflops"],
                    ms_per_launch=ms_launch, flops_per_launch=flops_launch,
                    hbm_algorithmic_gbs=hbm_alg / 1e9, hbm_peak_gbs=hbm_peak / 1e9, hbm_frac=hbm_alg / hbm_peak,
                    hbm_measured=dict(read_gbs=peaks["hbm_read_gbs"], copy_gbs=peaks["hbm_copy_gbs"],
                                      bytes=peaks["hbm_bytes"]))
    batched = None
    if not args.no_batched:
        batched = _leg("batched", batched_leg)
    adjoint = None
    if not args.no_adjoint:
        adjoint = _leg("adjoint", adjoint_leg)
    workloads = None
    if not args.no_workloads:
        workloads = _leg("workloads", workloads_leg)
    adaweno = None
    if not args.no_adaweno:
        adaweno = _leg("adaweno developed", adaweno_leg)
    driver = None
    if not args.no_driver:
        driver = _leg("fortran driver", driver_leg)
    cpu = None
    if not args.no_cpu:
        try:
            cpu = _leg("cpu baseline", cpu_leg, _threads_available(), args.cpu_steps)
        except Exception as e:  # the GPU number stands without the baseline
            cpu = dict(value=None, error=str(e))
    out = {
        "metric": "FP64 cell-updates/s, WENO-Z char. DMR; batched & adjoint vs plain GPU",
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": taken,
        "warmup": args.warmup,
        "ms_per_step": wall / taken * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (double Mach reflection initial condition)",
        "config": {"workload": "double Mach reflection 1920x480, characteristic WENO-Z + GLF (switch 104/255), "
                               "SSP-RK3, CFL 0.6, FP64", "ni": NI, "nj": NJ, "splitting": SPLITTING,
                   "recon": RECON, "cfl": CFL, "tile": "16x15"},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "batched": batched,
        "adjoint": adjoint,
        "other_workloads": workloads,
        "adaweno_developed": adaweno,
        "driver": driver,
    }
    if cpu and cpu.get("value"):
        out["speedup_vs_cpu"] = value / cpu["value"]
    print(json.dumps(out))


if __name__ == "__main__":
    main()
