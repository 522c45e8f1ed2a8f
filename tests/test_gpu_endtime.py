"""The configured workloads run as the reference runs them, against digests
of the reference's own runs (tests/golden/make_digests.py).

The reference's loop (general_main.f90:23-37) steps while current_time <
total_time with no clipping of the last step, so a run to t_end ends at the
first t >= t_end.  For each digest the GPU starts from the reference's initial
state (the oracle's init_state, pinned bitwise by the golden fixtures), runs
the same loop through dense_run and must give the same step count, the final
time and every step size the reference printed (to 1e-10 relative), and the
final state within the workloads' bar: the digest carries the density
averaged over 16x16-cell blocks, whose L1 difference is a lower bound of the
full field's and is held to 1e-9, and whose largest difference is held to
1e-11 of the largest block density.

Cases: double Mach 960x240 and 1920x480 to t = 0.2 (characteristic WENO-Z),
the reference's benchmark-1 vortex at 200x200 for 100 steps (LLF + WENO-JS),
Rayleigh-Taylor 1024x4096 for 20 steps and 256x1024 to t = 0.975 (AdaWENO
with gravity).  Rayleigh-Taylor 256x1024 run on to t = 1.95 reaches the
turbulent mixing layer, where any change of rounding grows exponentially
(measured: 3e-14 density L1 at t = 0.975, 7e-13 at t = 1.46, 6e-6 at t = 1.95
between two builds that differ only in FMA contraction): there the check is
the clock (the reference's step count to within 0.5 %, the first t >= 1.95)
and the block-averaged density to 1e-4 L1.
"""
import json
import os

import numpy as np
import pytest

from conftest import GOLDEN, dense

pytestmark = pytest.mark.gpu

with open(os.path.join(GOLDEN, "digests.json")) as _f:
    DIGESTS = json.load(_f)
CHAOTIC = {"rt_105_256x1024_t1.95"}


def _digest_mod():
    import sys
    if GOLDEN not in sys.path:
        sys.path.insert(0, GOLDEN)
    import digest
    return digest


@pytest.mark.parametrize("name", sorted(DIGESTS))
def test_run_matches_the_reference_digest(name, oracle):
    g = DIGESTS[name]
    z = np.load(os.path.join(GOLDEN, f"digest_{name}.npz"))
    dg = _digest_mod()
    D = dense()
    bm, ni, nj, sp, rc, cfl = g["benchmark"], g["ni"], g["nj"], g["splitting"], g["recon"], g["cfl"]
    o = oracle.Oracle(oracle.problem_params(bm, ni, nj, splitting=sp, recon=rc, cfl=cfl))
    o.init_state(bm)
    ic = o.cons()
    o.close()
    t_end = g["until_time"] if g["until_time"] is not None else float("inf")
    print(f"{name}: {ni}x{nj} benchmark {bm} {sp}/{rc}, reference {g['step']} steps to t = {g['time']!r}",
          flush=True)
    cfg = D.make_cfg(ni, nj, benchmark=bm, splitting=sp, recon=rc, cfl=cfl)
    with D.Solver(cfg) as s:
        s.set_state(ic)
        taken, dts = s.run(g["max_steps"], total_time=t_end, want_dts=True)
        got = s.interior()
        t, step = s.current_time, s.current_step
    coarse = dg.coarse_density(got)
    diff = np.abs(coarse - z["coarse"])
    block = dg.COARSE * dg.COARSE * cfg.dx * cfg.dy
    l1 = float(np.sum(diff) * block)
    rel = float(np.max(diff) / np.max(np.abs(z["coarse"])))
    print(f"{name}: GPU {step} steps to t = {t!r}; block-averaged density L1 {l1:.3e}, max rel {rel:.3e}",
          flush=True)
    assert np.all(np.isfinite(got))
    if g["until_time"] is not None:
        assert t >= t_end and t - dts[-1] < t_end  # the first t >= t_end, no clipping
    if name in CHAOTIC:
        assert abs(step - g["step"]) <= 0.005 * g["step"]
        assert l1 <= 1e-4
        return
    assert step == g["step"] and taken == g["step"]
    assert t == pytest.approx(g["time"], rel=1e-12, abs=0)
    np.testing.assert_allclose(dts, z["dts"], rtol=1e-10, atol=0)
    assert l1 <= 1e-9 and rel <= 1e-11
