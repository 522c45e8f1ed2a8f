"""The oracle (C restatement) against the reference's own outputs.

The golden fixtures were written by the reference solver itself (compiled
from its sources, tests/golden/make_golden.py).  The oracle must reproduce
them bit for bit, including the printed dt of every step.
"""
import os

import numpy as np
import pytest

from conftest import GOLDEN, golden, golden_cases

CASES = golden_cases()


@pytest.mark.parametrize("name", sorted(CASES))
def test_oracle_matches_reference_bitwise(name, oracle):
    c = CASES[name]
    g = golden(name)
    kw = dict(splitting=c["splitting"], recon=c["recon"], cfl=c["cfl"], stages=c.get("stages", 3),
              if_const_dt=int(c.get("if_const_dt", False)), dt_const=c.get("dt", 0.01))
    p = oracle.problem_params(c["benchmark"], c["ni"], c["nj"], **kw)
    o = oracle.Oracle(p)
    o.init_state(c["benchmark"])
    dts = []
    for _ in range(c["steps"]):
        o.step()
        dts.append(o.dt)
    got = o.interior()
    assert got.shape == g["cons"].shape
    assert np.array_equal(got, g["cons"]), f"max |diff| = {np.max(np.abs(got - g['cons']))}"
    assert o.time == float(g["time"])
    # the reference prints dt with 17 significant digits: exact round trip
    assert np.array_equal(np.array(dts), g["dts"])


def test_oracle_thread_count_invariance(oracle):
    """Row loops over OpenMP threads give the same bits (the only reduction is max)."""
    outs = []
    for nt in (1, 4):
        p = oracle.problem_params(5, 64, 16, splitting=104, recon=255, cfl=0.6, nthreads=nt)
        o = oracle.Oracle(p)
        o.init_state(5)
        o.step(5)
        outs.append(o.interior())
    assert np.array_equal(outs[0], outs[1])


DUMPS = sorted(f[5:-4] for f in os.listdir(GOLDEN) if f.startswith("dump_"))


@pytest.mark.parametrize("name", DUMPS)
def test_oracle_padded_state_matches_reference_dump(name, oracle):
    """The reference's whole .dat2d array at a save: interior, the ghost layers
    of its last boundary pass (the last stage's input at t^n) and zero
    corners.  The oracle's padded cons follows the same storage, bit for bit."""
    c = CASES[name]
    d = golden(f"dump_{name}")
    kw = dict(splitting=c["splitting"], recon=c["recon"], cfl=c["cfl"], stages=c.get("stages", 3),
              if_const_dt=int(c.get("if_const_dt", False)), dt_const=c.get("dt", 0.01))
    o = oracle.Oracle(oracle.problem_params(c["benchmark"], c["ni"], c["nj"], **kw))
    o.init_state(c["benchmark"])
    o.step(c["steps"])
    assert np.array_equal(o.cons(), d["cons"])
