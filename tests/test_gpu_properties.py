"""Size-independent properties of the GPU path at sizes the oracle is too slow
for, for every splitting (and each reconstruction of the component-wise ones):

* periodic shift equivariance: rolling the initial state by a number of cells
  that is not a multiple of any tile shape rolls the result bit for bit (the
  update of a cell depends only on its stencil and the global wave speeds, so
  tile seams, queue order and periodic ghost cells must not leave a trace);
* conservation: with periodic boundaries the fluxes telescope, so the totals
  of rho, rho u, rho v, E change only by rounding;
* free stream: a uniform state stays uniform (every cell bitwise equal to
  every other) and equal to the initial state to the rounding of the RK
  weights;
* re-read parameters: a solver switched mid-run by dense_update_params
  continues exactly like a fresh solver started from the same state.
"""
import numpy as np
import pytest

from conftest import dense

pytestmark = pytest.mark.gpu

SCHEMES = [(101, 251), (101, 253), (101, 255), (102, 253), (102, 255), (103, 255), (104, 255), (105, 255)]


def _periodic_cfg(ni, nj, h, sp, rc, cfl=0.5):
    D = dense()
    return D.make_cfg(ni, nj, benchmark=0, splitting=sp, recon=rc, cfl=cfl, gamma=1.4, dx=h, dy=h,
                      bc=[D.BC_PERIODIC] * 4)


def _rough_periodic_state(ni, nj, h, seed=0):
    """A vortex in a mean flow, a density band with a pressure jump (moving
    shocks and a contact) and seeded smooth noise: smooth and shocked regions,
    so that the AdaWENO branch takes both sides."""
    rng = np.random.default_rng(seed)
    x = (np.arange(ni) + 0.5) * h
    y = (np.arange(nj) + 0.5) * h
    X, Y = np.meshgrid(x, y)
    Lx, Ly = ni * h, nj * h
    dx = X - 0.3 * Lx
    dy = Y - 0.5 * Ly
    r2 = (dx * dx + dy * dy) / (0.05 * Lx) ** 2
    f = 0.8 * np.exp(0.5 * (1.0 - r2))
    u = 1.0 - f * dy / (0.05 * Lx)
    v = 0.5 + f * dx / (0.05 * Lx)
    band = (X > 0.6 * Lx) & (X < 0.8 * Lx)
    rho = np.where(band, 2.0, 1.0) + 0.01 * np.sin(2 * np.pi * (3 * X / Lx + 2 * Y / Ly) + rng.uniform(0, 6))
    p = np.where(band, 3.0, 1.0) * (1.0 + 0.01 * np.cos(2 * np.pi * 5 * Y / Ly))
    E = p / 0.4 + 0.5 * rho * (u * u + v * v)
    return np.stack([rho, rho * u, rho * v, E])


def _run(cfg, cons, steps):
    D = dense()
    with D.Solver(cfg) as s:
        s.set_state(cons)
        dts = np.array([s.step() for _ in range(steps)])
        return s.interior(), dts


@pytest.mark.parametrize("sp,rc", SCHEMES)
def test_periodic_shift_equivariance(sp, rc):
    ni, nj, h, steps = 640, 320, 1.0 / 320, 20
    shift = (45, 123)                     # rows, columns
    cons = _rough_periodic_state(ni, nj, h)
    cfg = _periodic_cfg(ni, nj, h, sp, rc)
    print(f"shift equivariance {sp}/{rc} {ni}x{nj} by {shift}", flush=True)
    a, da = _run(cfg, cons, steps)
    b, db = _run(cfg, np.roll(cons, shift, axis=(1, 2)), steps)
    assert np.all(np.isfinite(a))
    assert np.array_equal(da, db)
    assert np.array_equal(np.roll(a, shift, axis=(1, 2)), b)


@pytest.mark.parametrize("sp,rc", SCHEMES)
def test_periodic_conservation(sp, rc):
    ni, nj, h, steps = 1024, 512, 1.0 / 512, 100
    cons = _rough_periodic_state(ni, nj, h, seed=1)
    cfg = _periodic_cfg(ni, nj, h, sp, rc)
    print(f"conservation {sp}/{rc} {ni}x{nj} {steps} steps", flush=True)
    got, _ = _run(cfg, cons, steps)
    t0 = cons.reshape(4, -1).sum(axis=1)
    t1 = got.reshape(4, -1).sum(axis=1)
    rel = np.abs(t1 - t0) / np.abs(cons).reshape(4, -1).sum(axis=1)
    print(f"relative change of totals {rel}", flush=True)
    assert np.all(rel <= 1e-13)
    assert np.max(np.abs(got - cons)) > 1e-3          # the flow did move


@pytest.mark.parametrize("sp,rc", SCHEMES)
def test_free_stream_stays_uniform(sp, rc):
    ni, nj, steps = 1920, 480, 5
    prim = (1.4, 0.3, -0.2, 1.0)
    rho, u, v, p = prim
    q = np.array([rho, rho * u, rho * v, p / 0.4 + 0.5 * rho * (u * u + v * v)])
    cons = np.broadcast_to(q[:, None, None], (4, nj, ni)).copy()
    cfg = _periodic_cfg(ni, nj, 1.0 / nj, sp, rc)
    got, dts = _run(cfg, cons, steps)
    for k in range(4):
        assert np.all(got[k] == got[k, 0, 0]), f"component {k} not uniform"
    assert np.allclose(got[:, 0, 0], q, rtol=8 * np.finfo(float).eps, atol=0)
    assert np.all(dts == dts[0])


def test_update_params_switch_and_cfl_continue_like_fresh_solvers():
    """dense_update_params mid-run (the reference re-reads input.in at every
    save, general_main.f90:53): switching 104 -> 101 (another kernel on the
    same tile shape) and then changing the CFL number continues exactly as
    fresh solvers started from the intermediate states do."""
    D = dense()
    ni, nj, h = 320, 320, 1.0 / 320
    cons = _rough_periodic_state(ni, nj, h, seed=3)
    a = D.Solver(_periodic_cfg(ni, nj, h, 104, 255))
    try:
        a.set_state(cons)
        a.step(4)
        mid1 = a.interior()
        a.update_params(splitting=101, recon=255)
        a.step(4)
        mid2 = a.interior()
        a.update_params(cfl=0.3)
        dts_a = [a.step() for _ in range(4)]
        end = a.interior()
    finally:
        a.close()
    ref2, _ = _run(_periodic_cfg(ni, nj, h, 101, 255), mid1, 4)
    assert np.array_equal(mid2, ref2)
    ref3, dts_ref = _run(_periodic_cfg(ni, nj, h, 101, 255, cfl=0.3), mid2, 4)
    assert np.array_equal(end, ref3)
    assert np.array_equal(np.array(dts_a), dts_ref)


def test_adaweno_threshold_one_equals_characteristic_at_960x240():
    """theta <= 1, so AdaWENO with delta = 1 never takes the component-wise
    branch (scheme_weno5th_SC.f90:110) and equals switch 104 bit for bit; at
    the default delta = 0.5 it differs.  Double Mach 960x240, 20 steps."""
    D = dense()
    bench = __import__("bench")
    ic = bench.dmr_initial_state(960, 240)
    outs = {}
    for sp, delta in ((104, 0.5), (105, 1.0), (105, 0.5)):
        cfg = D.make_cfg(960, 240, benchmark=5, splitting=sp, recon=255, cfl=0.6, delta=delta)
        outs[(sp, delta)], _ = _run(cfg, ic, 20)
    assert np.array_equal(outs[(105, 1.0)], outs[(104, 0.5)])
    assert not np.array_equal(outs[(105, 0.5)], outs[(104, 0.5)])
