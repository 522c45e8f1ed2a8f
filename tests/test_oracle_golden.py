"""The oracle (C restatement) against the reference's own outputs.

The golden fixtures were written by the reference solver itself (compiled
from its sources, tests/golden/make_golden.py).  The oracle must reproduce
them bit for bit, including the printed dt of every step.
"""
import numpy as np
import pytest

from conftest import golden, golden_cases

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
