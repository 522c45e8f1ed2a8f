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


@pytest.mark.parametrize("splitting,recon", [(101, 255), (104, 255), (105, 255)])
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
