"""The stepping entry points' contracts through the C ABI.

* dense_run returns the dt of every step taken, however many (the device
  history grows to max_steps), and replays exactly max_steps steps for any
  count (16-step graphs plus one graph of the remainder), bit for bit the
  steps dense_step takes one at a time.
* dense_run_saving also stops at the reference's save clock: elasp_time is
  advanced by dt after every step (general_main.f90:37) and the run ends after
  the first step at which elasp_time >= dt_save (general_main.f90:40).
* dense_get_dump returns the array the reference writes at a save
  (read_WriteAndLoad.f90:1-28): ghost layers of its last boundary pass, the
  last stage's input at t^n (general_main.f90:26-31), and zero corners --
  compared with whole dumps the reference wrote (tests/golden/dump_*.npz) to
  the workloads' tolerances, corners exactly.
* dense_autotune's trials overwrite the stage buffers, so theta and the dump
  ghosts of the last step are refused or recomputed until the next step.
"""
import importlib
import os

import numpy as np
import pytest

from conftest import GOLDEN, assert_parity, dense, golden, golden_cases

pytestmark = pytest.mark.gpu
CASES = golden_cases()
DUMPS = sorted(f[5:-4] for f in os.listdir(GOLDEN) if f.startswith("dump_"))


def _vortex(n=24, splitting=101, recon=251, **kw):
    D = dense()
    B = importlib.import_module("2dense_amd.batch")
    cons, h = B.vortex_state(n, 4.0, 3.0, 6.0)
    cfg = D.make_cfg(n, n, benchmark=0, splitting=splitting, recon=recon, cfl=0.5, gamma=1.4, dx=h, dy=h,
                     bc=[D.BC_PERIODIC] * 4, **kw)
    s = D.Solver(cfg)
    s.set_state(cons)
    return s


def _case_solver(name, oracle):
    D = dense()
    c = CASES[name]
    p = oracle.problem_params(c["benchmark"], c["ni"], c["nj"], splitting=c["splitting"], recon=c["recon"],
                              cfl=c["cfl"])
    o = oracle.Oracle(p)
    o.init_state(c["benchmark"])
    cfg = D.make_cfg(c["ni"], c["nj"], benchmark=c["benchmark"], splitting=c["splitting"], recon=c["recon"],
                     cfl=c["cfl"], stages=c.get("stages", 3), if_const_dt=c.get("if_const_dt", False),
                     dt=c.get("dt", 0.01))
    s = D.Solver(cfg)
    s.set_state(o.cons())
    return s


def test_dt_history_holds_every_step_of_a_long_run():
    """5,000 steps in one dense_run: every dt is returned (the history used to
    stop at 4,096 entries) and equals the dt of the same steps taken one by one."""
    n = 5000
    print("dt history: 24x24 vortex, 101/251, 5000 steps", flush=True)
    with _vortex() as a, _vortex() as b:
        taken, dts = a.run(n, want_dts=True)
        assert taken == n and dts.shape == (n,)
        one = np.array([b.step() for _ in range(n)])
        assert np.array_equal(dts, one)
        assert np.array_equal(a.interior(), b.interior())
        assert a.current_time == b.current_time


def test_dt_history_of_the_host_stepped_inflow(oracle):
    """The shear-layer inflow steps from the host; its history fills the
    caller's buffer of max_steps entries (5,000 steps, more than 4,096)."""
    print("dt history: shear layer 64x32, 101/255, 5000 steps", flush=True)
    with _case_solver("ssl_101_255", oracle) as s:
        taken, dts = s.run(5000, want_dts=True)
        assert taken == 5000 and dts.shape == (5000,) and np.all(dts > 0)
        t = 0.0
        for dt in dts:  # current_time = current_time + dt, general_main.f90:36
            t = t + float(dt)
        assert s.current_time == t


@pytest.mark.parametrize("counts", [(1, 5, 16), (17, 33), (15, 2, 40)])
def test_run_of_any_step_count_equals_single_steps(counts):
    """Consecutive runs of counts that are and are not multiples of the graph
    chunk: exactly the requested steps, each stage launch a working one, the
    state and clock those of single steps."""
    with _vortex(splitting=104) as a, _vortex(splitting=104) as b:
        for k in counts:
            taken = a.run(k)
            ms, launches = a.last_run_gpu_ms()
            assert taken == k and launches == 3 * k
            b.step(k)
            assert (a.current_step, a.current_time) == (b.current_step, b.current_time)
        assert np.array_equal(a.interior(), b.interior())


def test_run_saving_stops_at_the_save_clock():
    with _vortex() as a, _vortex() as b:
        one = np.array([b.step() for _ in range(40)])
        # the reference's clocks, replayed on the host
        el, stop = 0.0, None
        dt_save = float(np.sum(one[:7]) - 0.5 * one[6])  # reached during the 7th step
        for k, dt in enumerate(one):
            el = el + dt
            if el >= dt_save:
                stop = k + 1
                break
        assert stop == 7
        taken, dts, el_dev = a.run_saving(40, dt_save=dt_save, elasp=0.0)
        assert taken == stop and np.array_equal(dts, one[:stop]) and el_dev == el
        # elasp_time already past dt_save (no Tecplot save resets it): the
        # reference saves after every step, so one step runs
        taken, dts, el2 = a.run_saving(40, dt_save=dt_save, elasp=el_dev)
        assert taken == 1 and dts[0] == one[stop] and el2 == el_dev + one[stop]
        # a far save clock: the step limit ends the run
        taken, dts, _ = a.run_saving(9, dt_save=1e30, elasp=0.0)
        assert taken == 9 and np.array_equal(dts, one[stop + 1:stop + 10])


@pytest.mark.parametrize("name", DUMPS)
def test_dump_equals_the_reference_dump(name, oracle):
    """The whole .dat2d array, ghost layers and corners included."""
    c = CASES[name]
    d = golden(f"dump_{name}")
    print(f"dump {name}", flush=True)
    with _case_solver(name, oracle) as s:
        for _ in range(c["steps"]):
            s.step()
        got = s.get_dump()
        assert_parity(got, d["cons"], s.cfg.dx, s.cfg.dy, f"dump {name}")
        corners = np.zeros(got.shape[1:], dtype=bool)
        corners[:4, :4] = corners[:4, -4:] = corners[-4:, :4] = corners[-4:, -4:] = True
        assert np.all(got[:, corners] == 0.0) and np.all(d["cons"][:, corners] == 0.0)
        # the ghost layers are those of the last stage's input: a boundary pass
        # over the final interior (dense_get_state) fills them differently
        state = s.get_state()
        assert np.array_equal(state[:, 4:-4, 4:-4], got[:, 4:-4, 4:-4])
        assert not np.array_equal(state, got)


def test_dump_before_any_step_and_one_stage_runs():
    D = dense()
    with _vortex() as s:
        assert np.array_equal(s.get_dump(), s.get_state())
    with _vortex(stages=1) as s:
        s.step()
        with pytest.raises(D.DenseError) as e:
            s.get_dump()
        assert e.value.code == -2


def test_autotune_invalidates_theta_and_dump_ghosts():
    D = dense()
    with _vortex(n=48, splitting=105, recon=255) as s:
        s.step(2)
        s.theta()
        s.autotune(trial_steps=2)
        with pytest.raises(D.DenseError):
            s.theta()
        assert np.array_equal(s.get_dump(), s.get_state())
        s.step()
        s.theta()
