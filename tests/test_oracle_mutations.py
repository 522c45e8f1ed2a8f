"""The reference's fixtures pin the oracle: each plausible mistake, planted in
the restatement one at a time (g_mut in oracle/dense_oracle.c), makes at least
one fixture written by the reference fail, while the restatement as written
matches them all (tests/test_oracle_golden.py).  The mistakes cover a dropped
single-precision literal (the WENO epsilon, the stage-3 body-force
coefficient), a wrong wave speed (global instead of per-cell Lax-Friedrichs),
the clock of a time-dependent boundary, the evaluation order of an
extrapolation and of the CFL step, transposed operands (Roe weights), an RK
weight one ulp off, and the characteristic schemes following the
reconstruction switch (the reference always uses WENO-Z there)."""
import numpy as np
import pytest

from conftest import golden, golden_cases

CASES = golden_cases()

# mutation -> fixtures that must catch it
MUTATIONS = {
    1: ("WENO epsilon as the double 1e-6", ["ivc_101_253", "dmr_104"]),
    2: ("stage-3 body-force coefficient 2/3 in double", ["rt_105"]),
    3: ("LLF with the global wave speed", ["ivc_101_255", "sv_101_255"]),
    4: ("double-Mach top boundary at t^n + dt", ["dmr_104"]),
    5: ("extrapolation as (4 q1 - q2)/3", ["ivc_101_251", "dmr_105"]),
    6: ("Roe weights swapped", ["ivc_104", "rm_105"]),
    7: ("stage-3 weight 1 - 1/3", ["ivc_101_253", "sedov_104"]),
    8: ("characteristic schemes follow the reconstruction switch", ["dmr_104_recon251", "sv_105_recon253"]),
    9: ("CFL step with cfl*(dt_xi*dt_eta) associated", ["dmr_101_253"]),
}


def _matches(oracle, name):
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
    ok = np.array_equal(o.interior(), g["cons"]) and np.array_equal(np.array(dts), g["dts"])
    o.close()
    return ok


@pytest.mark.parametrize("m", sorted(MUTATIONS))
def test_planted_mistake_is_caught_by_reference_fixtures(m, oracle):
    what, fixtures = MUTATIONS[m]
    for name in fixtures:
        assert _matches(oracle, name), f"{name} must match without the planted mistake"
    with oracle.planted_mistake(m):
        caught = [name for name in fixtures if not _matches(oracle, name)]
    assert caught == fixtures, f"'{what}' not caught by {sorted(set(fixtures) - set(caught))}"
    for name in fixtures:
        assert _matches(oracle, name), "the planted mistake must be removed again"
