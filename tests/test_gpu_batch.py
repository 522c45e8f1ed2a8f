"""Batched mode: every instance of an ensemble advanced by one launch per RK
stage equals the same instance advanced alone (bit for bit), for every
schedule the autotuner can choose."""
import importlib

import numpy as np
import pytest

from conftest import dense

pytestmark = pytest.mark.gpu


def _batch_mod():
    return importlib.import_module("2dense_amd.batch")


def _ensemble(n=6, seed=7, splitting=101, recon=255):
    B = _batch_mod()
    return B.make_ensemble(n=n, seed=seed, nmin=16, nmax=72, splitting=splitting, recon=recon)


def _alone(cfg, cons, steps):
    D = dense()
    with D.Solver(cfg) as s:
        s.set_state(cons)
        dts = [s.step() for _ in range(steps)]
        return s.interior(), np.array(dts)


@pytest.mark.parametrize("splitting,recon", [(101, 255), (103, 253), (104, 255), (105, 255)])
def test_batch_equals_single_instances(splitting, recon):
    B = _batch_mod()
    cfgs, states, insts = _ensemble(splitting=splitting, recon=recon)
    steps = 8
    print(f"batch {splitting}/{recon}: sizes {[(i.ni, i.nj) for i in insts]}", flush=True)
    with B.Batch(cfgs) as b:
        for i, c in enumerate(states):
            b.set_state(i, c)
        hist = b.run(steps, want_dts=True)
        for i in range(len(cfgs)):
            ref, dts = _alone(cfgs[i], states[i], steps)
            got = b.interior(i)
            assert np.array_equal(got, ref), f"instance {i} ({insts[i].kind}) max |diff| {np.max(np.abs(got - ref))}"
            assert np.array_equal(hist[:, i], dts), f"instance {i} dt sequence"
            assert b.t[i] == pytest.approx(dts.sum(), rel=0, abs=0) or b.t[i] == np.cumsum(dts)[-1]


def test_batch_dt_history_covers_every_step():
    """dense_batch_run returns every step's dt per instance for any run length
    (the device history grows past its initial 256 steps), equal to the same
    instances run alone; a second, shorter run reuses it."""
    B = _batch_mod()
    D = dense()
    cfgs, states, _ = _ensemble(n=2, seed=3)
    with B.Batch(cfgs) as b:
        for i, c in enumerate(states):
            b.set_state(i, c)
        hist = b.run(300, want_dts=True)
        hist2 = b.run(40, want_dts=True)
    assert hist.shape == (300, 2) and hist2.shape == (40, 2)
    for i in range(2):
        with D.Solver(cfgs[i]) as s:
            s.set_state(states[i])
            n, dts = s.run(340, want_dts=True)
        assert n == 340
        assert np.array_equal(hist[:, i], dts[:300]), f"instance {i}: first run"
        assert np.array_equal(hist2[:, i], dts[300:]), f"instance {i}: second run"


@pytest.mark.parametrize("tile", [0, 1, 2, 3])
@pytest.mark.parametrize("order,grid", [(0, 100), (1, 50)])
def test_batch_schedule_invariance(tile, order, grid):
    B = _batch_mod()
    cfgs, states, _ = _ensemble(n=5, seed=11)
    outs = []
    for sched in [(0, 0, 100), (tile, order, grid)]:
        with B.Batch(cfgs) as b:
            b.set_schedule(*sched)
            for i, c in enumerate(states):
                b.set_state(i, c)
            b.run(6)
            outs.append([b.interior(i) for i in range(len(cfgs))])
    for a, c in zip(*outs):
        assert np.array_equal(a, c)


@pytest.mark.parametrize("splitting", [101, 102])
@pytest.mark.parametrize("tile", [0, 3])
def test_batch_upwind5_equals_single_instances(splitting, tile):
    """Lax-Friedrichs + upwind5 ensembles, including the 32x15 tiles whose face
    fluxes overwrite the cell records in LDS, equal each instance alone."""
    B = _batch_mod()
    cfgs, states, insts = _ensemble(n=5, seed=5, splitting=splitting, recon=251)
    steps = 6
    print(f"batch {splitting}/251 tile {tile}: sizes {[(i.ni, i.nj) for i in insts]}", flush=True)
    with B.Batch(cfgs) as b:
        b.set_schedule(tile, 0, 100)
        for i, c in enumerate(states):
            b.set_state(i, c)
        b.run(steps)
        for i in range(len(cfgs)):
            ref, _ = _alone(cfgs[i], states[i], steps)
            assert np.array_equal(b.interior(i), ref), f"instance {i} ({insts[i].kind})"


def test_batch_autotune_restores_state():
    B = _batch_mod()
    cfgs, states, _ = _ensemble(n=4, seed=3)
    with B.Batch(cfgs) as b:
        for i, c in enumerate(states):
            b.set_state(i, c)
        tile, order, grid, ms = b.autotune(trial_steps=2)
        assert 0 <= tile <= 3 and order in (0, 1) and 1 <= grid <= 100 and ms > 0
        b.run(5)
        got = [b.interior(i) for i in range(len(cfgs))]
    for i in range(len(cfgs)):
        ref, _ = _alone(cfgs[i], states[i], 5)
        assert np.array_equal(got[i], ref)


def test_batch_rejects_shear_layer_inflow():
    """The shear-layer inflow needs the host's cos of t^n at every step: not batched."""
    D = dense()
    B = _batch_mod()
    cfg = D.make_cfg(64, 32, benchmark=6, splitting=D.LLF, recon=D.WENOZ)
    with pytest.raises(D.DenseError) as e:
        B.Batch([cfg])
    assert e.value.code == -2


def _centres_x(cfg):
    return (np.arange(cfg.ni) + 0.5) * cfg.dx


def test_device_initial_states_match_host_builders():
    """The device generator evaluates the host builders' formulas in the same
    order; only exp and pow may round differently, so every component agrees
    to a few units in the last place of its largest value, and the
    Rankine-Hugoniot region (no exp/pow) agrees bit for bit."""
    B = _batch_mod()
    cfgs, states, insts = B.make_ensemble(n=10, seed=3, nmin=16, nmax=96)
    print(f"device ICs: sizes {[(i.ni, i.nj) for i in insts]}", flush=True)
    with B.Batch(cfgs) as b:
        b.init_ic(insts)
        assert np.all(b.t == 0.0)
        for i, inst in enumerate(insts):
            got, ref = b.interior(i), states[i]
            assert got.shape == ref.shape and np.all(np.isfinite(got))
            for k in range(4):
                err = np.max(np.abs(got[k] - ref[k])) / np.max(np.abs(ref[k]))
                assert err <= 1e-14, f"instance {i} ({inst.kind}) component {k}: {err:.3e}"
            if inst.kind == "shock_vortex":
                right = _centres_x(cfgs[i]) >= 0.5
                assert right.any()
                assert np.array_equal(got[:, :, right], ref[:, :, right])


def test_device_initial_states_evolve_like_host_states():
    B = _batch_mod()
    cfgs, states, insts = B.make_ensemble(n=6, seed=9, nmin=16, nmax=64)
    steps = 30
    with B.Batch(cfgs) as dev, B.Batch(cfgs) as host:
        dev.init_ic(insts)
        for i, c in enumerate(states):
            host.set_state(i, c)
        hd = dev.run(steps, want_dts=True)
        hh = host.run(steps, want_dts=True)
        assert np.allclose(hd, hh, rtol=1e-12, atol=0)
        for i in range(len(cfgs)):
            a, r = dev.interior(i), host.interior(i)
            err = np.max(np.abs(a - r)) / np.max(np.abs(r))
            assert err <= 1e-11, f"instance {i} ({insts[i].kind}): {err:.3e}"


def test_device_initial_states_reject_bad_descriptors():
    D = dense()
    B = _batch_mod()
    cfgs, _, insts = B.make_ensemble(n=2, seed=1, nmin=16, nmax=24, host_states=False)
    with B.Batch(cfgs) as b:
        good = [B.ic_desc(x) for x in insts]
        for field, value in (("kind", 7), ("mach", 0.5), ("rc", 0.0), ("eps", float("nan"))):
            bad = [B.ic_desc(x) for x in insts]
            k = 1 if field in ("mach", "rc") else 0
            setattr(bad[k], field, value)
            with pytest.raises(D.DenseError) as e:
                b.init_ic(bad)
            assert e.value.code == -1
        b.init_ic(good)
        b.run(2)
        assert np.all(np.isfinite(b.interior(0)))


@pytest.mark.parametrize("splitting,recon", [(101, 251), (101, 255), (104, 255)])
def test_batch_with_large_instances_on_512_thread_tiles(splitting, recon):
    """An ensemble whose largest instances (512^2 and a 512x256 shock/vortex)
    hold far more tiles than the persistent grid has workgroups, on the 32x15
    schedule with 512 threads: each workgroup pulls many tiles across
    instance boundaries, and (for upwind5) writes its face fluxes over the
    cell records.  Every instance equals its single-block run bit for bit."""
    D = dense()
    B = _batch_mod()
    big_v = B.vortex_state(512, 3.5, 2.0, 7.0)
    big_sv = B.shock_vortex_state(512, 256, 1.3, 0.25, 0.4)
    cfgs, states, _ = _ensemble(n=4, seed=9, splitting=splitting, recon=recon)
    cfgs = list(cfgs) + [
        D.make_cfg(512, 512, benchmark=0, splitting=splitting, recon=recon, cfl=0.5, gamma=1.4, dx=big_v[1],
                   dy=big_v[1], bc=[D.BC_PERIODIC] * 4),
        D.make_cfg(512, 256, benchmark=0, splitting=splitting, recon=recon, cfl=0.5, gamma=1.4, dx=big_sv[1],
                   dy=big_sv[2], bc=[D.BC_EXTRAP, D.BC_EXTRAP, D.BC_REFLECT, D.BC_REFLECT])]
    states = list(states) + [big_v[0], big_sv[0]]
    steps = 5
    print(f"batch {splitting}/{recon} with 512^2 and 512x256 instances, tile 3", flush=True)
    with B.Batch(cfgs) as b:
        b.set_schedule(3, 1, 100)  # 32x15 with 512 threads, largest instance first
        for i, c in enumerate(states):
            b.set_state(i, c)
        b.run(steps)
        for i in range(len(cfgs)):
            ref, _ = _alone(cfgs[i], states[i], steps)
            assert np.array_equal(b.interior(i), ref), f"instance {i} ({cfgs[i].ni}x{cfgs[i].nj})"
