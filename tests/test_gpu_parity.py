"""HIP path vs the reference outputs and the oracle.

The GPU evaluates the reference's formulas with FMA contraction and
reciprocal-based division, so it is held to the workloads' tolerances
(conftest.assert_parity: relative L-inf <= 1e-11 per conserved component and
density L1 <= 1e-9; step sizes to 1e-12 relative) against the fixtures the
reference solver wrote, and to bit-for-bit equality within the build (tile
shapes, graph replay vs single steps).
"""
import numpy as np
import pytest

from conftest import assert_dts, assert_parity, dense, golden, golden_cases

pytestmark = pytest.mark.gpu
CASES = golden_cases()


def _oracle_ic(oracle, c):
    p = oracle.problem_params(c["benchmark"], c["ni"], c["nj"], splitting=c["splitting"], recon=c["recon"],
                              cfl=c["cfl"])
    o = oracle.Oracle(p)
    o.init_state(c["benchmark"])
    return o


def _solver(c, tile=0):
    D = dense()
    cfg = D.make_cfg(c["ni"], c["nj"], benchmark=c["benchmark"], splitting=c["splitting"], recon=c["recon"],
                     cfl=c["cfl"], stages=c.get("stages", 3), if_const_dt=c.get("if_const_dt", False),
                     dt=c.get("dt", 0.01), tile=tile)
    return D.Solver(cfg)


@pytest.mark.parametrize("name", sorted(CASES))
def test_gpu_matches_reference_golden(name, oracle):
    c = CASES[name]
    g = golden(name)
    o = _oracle_ic(oracle, c)
    print(f"case {name}", flush=True)
    with _solver(c) as s:
        s.set_state(o.cons())
        dts = []
        for _ in range(c["steps"]):
            dts.append(s.step())
        got = s.interior()
        dx, dy = s.cfg.dx, s.cfg.dy
    rel, l1 = assert_parity(got, g["cons"], dx, dy, name)
    print(f"{name}: relative L-inf {rel:.2e}, density L1 {l1:.2e}", flush=True)
    assert_dts(dts, g["dts"], name)
    assert s.current_time == pytest.approx(float(g["time"]), rel=1e-12, abs=0)


@pytest.mark.parametrize("name", ["dmr_104", "ivc_101_253", "rt_105", "sv_101_255", "ivc_101_251", "ivc_102_251",
                                  "dmr_103_251", "ivc_103_255"])
def test_gpu_tile_shape_invariance(name, oracle):
    """Every tile shape gives the same bits (faces are independent; the only
    cross-tile reduction is a max), and the reference's state within the
    tolerances.  The upwind5 cases cover the 512-thread tiles whose face
    fluxes overwrite the cell records in LDS; the Steger-Warming cases the
    split-flux bands of every shape."""
    c = CASES[name]
    g = golden(name)
    o = _oracle_ic(oracle, c)
    out = {}
    for tile in (0, 1, 2, 3, 4, 5):
        print(f"case {name} tile {tile}", flush=True)
        with _solver(c, tile=tile) as s:
            s.set_state(o.cons())
            s.step(c["steps"])
            out[tile] = s.interior()
            dx, dy = s.cfg.dx, s.cfg.dy
    assert_parity(out[0], g["cons"], dx, dy, name)
    for tile in (1, 2, 3, 4, 5):
        assert np.array_equal(out[tile], out[0]), f"tile {tile}"


@pytest.mark.parametrize("name", ["dmr_104", "ssl_104"])
def test_gpu_run_graph_matches_step(oracle, name):
    """dense_run (graph replay with the device clock; host-stepped for the
    shear-layer inflow) equals dense_step one by one."""
    c = CASES[name]
    o = _oracle_ic(oracle, c)
    with _solver(c) as a, _solver(c) as b:
        a.set_state(o.cons())
        b.set_state(o.cons())
        a.step(12)
        n, dts = b.run(12, want_dts=True)
        assert n == 12
        assert np.array_equal(a.interior(), b.interior())
        assert a.current_time == b.current_time


@pytest.mark.parametrize("name,total", [("dmr_104", 0.005), ("ssl_104", 3.0)])
def test_gpu_run_stops_at_total_time(oracle, name, total):
    """Loop condition current_time < total_time (general_main.f90:23)."""
    c = CASES[name]
    o = _oracle_ic(oracle, c)
    o.run(total, 10 ** 6)
    with _solver(c) as s:
        s.set_state(_oracle_ic(oracle, c).cons())
        s.run(1000, total_time=total)
        assert s.current_step == o.current_step
        assert s.current_time == pytest.approx(o.time, rel=1e-12, abs=0) and s.current_time >= total
        assert_parity(s.interior(), o.interior(), s.cfg.dx, s.cfg.dy, name)


@pytest.mark.parametrize("sp,rc", [(101, 251), (102, 251), (103, 255), (104, 255), (105, 255)])
def test_tile_shapes_agree_when_workgroups_take_many_tiles(sp, rc):
    """Double Mach 1920x480: 3,840 tiles of 16x15 (1,920 of 32x15, 17,280 of
    8x7) against the few hundred workgroups the persistent grid holds, so each
    workgroup pulls many tiles from its XCD's queue and rewrites its LDS for
    each: the 512-thread shapes whose face fluxes overwrite the cell records
    (32x15 and 16x15/512 with upwind5), the x/y halves of 16x15/512, the
    Steger-Warming bands and the characteristic faces all give the bits of
    the 16x15 shape after 4 steps."""
    D = dense()
    bench = __import__("bench")
    cons = bench.dmr_initial_state(1920, 480)
    out = {}
    for tile in (4, 1, 2, 3, 5):
        print(f"DMR 1920x480 {sp}/{rc} tile {tile}", flush=True)
        cfg = D.make_cfg(1920, 480, benchmark=5, splitting=sp, recon=rc, cfl=0.6, tile=tile)
        with D.Solver(cfg) as s:
            s.set_state(cons)
            s.run(4)
            out[tile] = s.interior()
    for tile in (1, 2, 3, 5):
        assert np.array_equal(out[tile], out[4]), f"tile {tile}: max |diff| {np.max(np.abs(out[tile] - out[4]))}"
