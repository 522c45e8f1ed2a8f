"""Order of accuracy on a smooth flow with an exact solution, in space and in
time.

The isentropic vortex in the periodic free stream u = v = 1 is transported
without change: at time T the exact state is the initial one translated by
(T, T).  With the step size tied to h^(5/3) (constant dt, so the RK3 error
O(dt^3) stays below the spatial O(h^5)), the density L1 error of every scheme
falls at the rate of its 5th-order stencils (weno_mod.f90,
scheme_upwind5_general.f90); a wrong stencil coefficient or weight would show
as a lower order.  Measured on MI355X, finest pair 160 -> 320: 5.2 (101/255),
4.7 (101/253), 5.0 (102/251), 4.9 (104), 4.9 (105), 5.3 (103/255).
"""
import importlib
import math

import numpy as np
import pytest

from conftest import dense

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("sp,rc", [(101, 255), (101, 253), (102, 251), (103, 255), (104, 255), (105, 255)])
def test_fifth_order_on_the_isentropic_vortex(sp, rc):
    D = dense()
    B = importlib.import_module("2dense_amd.batch")
    T = 2.0
    errs = []
    for n in (40, 80, 160, 320):
        cons, h = B.vortex_state(n, 5.0, 5.0, 5.0)
        steps = int(round(T / (0.05 * h ** (5.0 / 3.0))))
        cfg = D.make_cfg(n, n, benchmark=0, splitting=sp, recon=rc, gamma=1.4, dx=h, dy=h,
                         bc=[D.BC_PERIODIC] * 4, if_const_dt=True, dt=T / steps)
        with D.Solver(cfg) as s:
            s.set_state(cons)
            assert s.run(steps) == steps
            got = s.interior()
        exact, _ = B.vortex_state(n, 5.0, 5.0 + T, 5.0 + T)
        errs.append(float(np.sum(np.abs(got[0] - exact[0])) * h * h))
    orders = [math.log2(errs[i] / errs[i + 1]) for i in range(len(errs) - 1)]
    print(f"{sp}/{rc}: L1 {['%.3e' % e for e in errs]}, orders {['%.2f' % o for o in orders]}", flush=True)
    assert all(o > 3.5 for o in orders)
    assert orders[-1] > 4.4


@pytest.mark.parametrize("sp,rc", [(101, 255), (104, 255)])
def test_rk3_third_order_in_time(sp, rc):
    """SSP-RK3 (time_advance_general.f90:35-217) at a fixed grid: halving a
    constant step twice, the differences between successive solutions shrink by
    2^3 (a wrong stage weight would leave first or second order)."""
    D = dense()
    B = importlib.import_module("2dense_amd.batch")
    n, T = 160, 1.0
    cons, h = B.vortex_state(n, 5.0, 5.0, 5.0)
    sols = []
    for steps in (80, 160, 320):  # CFL 0.7, 0.35, 0.18
        cfg = D.make_cfg(n, n, benchmark=0, splitting=sp, recon=rc, gamma=1.4, dx=h, dy=h,
                         bc=[D.BC_PERIODIC] * 4, if_const_dt=True, dt=T / steps)
        with D.Solver(cfg) as s:
            s.set_state(cons)
            assert s.run(steps) == steps
            sols.append(s.interior())
    d1 = np.sum(np.abs(sols[0] - sols[1]))
    d2 = np.sum(np.abs(sols[1] - sols[2]))
    order = math.log2(d1 / d2)
    print(f"{sp}/{rc}: time order {order:.2f}", flush=True)
    assert 2.6 < order < 3.4
