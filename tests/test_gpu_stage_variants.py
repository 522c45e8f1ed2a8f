"""Every splitting and reconstruction with 1-, 2- and 3-stage steps and with a
constant step size, against the C oracle.

The reference's stage loop (general_main.f90:26-31) runs `stages` RK stages
with the stage formulas of time_advance_general.f90 (:35-54 stage 1,
:113-135 stage 2, :195-217 stage 3) and, with if_const_dt, the step `dt`
instead of the CFL step (:20-24).  The golden fixtures hold the 1- and
2-stage updates and the constant step for a few switch pairs; here every
pair runs each variant on a problem whose boundaries exercise the stage's
boundary time (double Mach's moving shock) or the body force
(Rayleigh-Taylor), 6 steps, with the step sizes.  The smooth problems are held
to the workloads' bars (relative L-inf <= 1e-11, density L1 <= 1e-9); the
double Mach reflection on 40x16 cells, whose Mach-10 shock spans a few cells,
to the bar the workloads state for shocked flows, density L1 <= 1e-9, with
relative L-inf <= 1e-9 beside it: there Steger-Warming + WENO-Z turns the
last-bit differences of FMA contraction and reciprocal division into 8e-11
(the density L1 is 1.5e-11).
"""
import numpy as np
import pytest

from conftest import L1_RHO, assert_dts, assert_parity, dense, parity_errors

pytestmark = pytest.mark.gpu

PAIRS = [(101, 251), (101, 255), (102, 253), (103, 255), (104, 255), (105, 255)]
VARIANTS = [  # (benchmark, ni, nj, stages, const dt)
    (5, 40, 16, 1, None),
    (5, 40, 16, 2, None),
    (3, 16, 48, 2, None),
    (3, 16, 48, 1, None),
    (1, 24, 24, 3, 0.02),
    (5, 40, 16, 3, 2.0e-3),
]


@pytest.mark.parametrize("variant", VARIANTS, ids=[f"b{v[0]}-s{v[3]}-{'const' if v[4] else 'cfl'}" for v in VARIANTS])
@pytest.mark.parametrize("splitting,recon", PAIRS)
def test_stage_counts_and_constant_step_match_the_oracle(oracle, splitting, recon, variant):
    benchmark, ni, nj, stages, dt_const = variant
    D = dense()
    cfl, steps = (0.3 if stages == 1 else 0.45), 6
    p = oracle.problem_params(benchmark, ni, nj, splitting=splitting, recon=recon, cfl=cfl, stages=stages,
                              if_const_dt=int(dt_const is not None), dt_const=dt_const or 0.01)
    o = oracle.Oracle(p)
    o.init_state(benchmark)
    cons0 = o.cons()
    dts = []
    for _ in range(steps):
        o.step()
        dts.append(o.dt)
    want = o.interior()
    o.close()
    what = f"benchmark {benchmark}, {splitting}/{recon}, stages {stages}, dt {dt_const or 'CFL'}"
    assert np.all(np.isfinite(want)), f"{what}: the oracle's run is not finite"
    cfg = D.make_cfg(ni, nj, benchmark=benchmark, splitting=splitting, recon=recon, cfl=cfl, stages=stages,
                     if_const_dt=dt_const is not None, dt=dt_const or 0.01)
    with D.Solver(cfg) as s:
        s.set_state(cons0)
        n, got_dts = s.run(steps, want_dts=True)
        got = s.interior()
    print(what, flush=True)
    assert n == steps
    if benchmark == 5:
        assert np.all(np.isfinite(got)), f"{what}: non-finite values"
        rel, l1 = parity_errors(got, want, cfg.dx, cfg.dy)
        assert l1 <= L1_RHO and rel <= 1e-9, f"{what}: relative L-inf {rel:.3e}, density L1 {l1:.3e}"
    else:
        assert_parity(got, want, cfg.dx, cfg.dy, what)
    assert_dts(got_dts, dts, what)
