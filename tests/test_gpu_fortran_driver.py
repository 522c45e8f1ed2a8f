"""The Fortran driver (iso_c_binding over the C ABI) end to end.

The driver reads the reference's input.in, builds the canned problem, steps
it on the GPU and writes 2d_<step>.dat2d in the reference's unformatted
layout; its files hold the reference's records (the same header records,
cell centres and metrics, bit for bit) and its state and printed dt sequence
match the reference's own outputs to the workloads' tolerances
(conftest.assert_parity).  Runs that differ only in what the driver does
around the GPU (restart from a dump, autotuned tile shape) are bit for bit.
"""
import os
import re
import subprocess

import numpy as np
import pytest

from conftest import ROOT, assert_dts, assert_parity, golden, golden_cases

pytestmark = pytest.mark.gpu
DRIVER = os.path.join(ROOT, "2dense_amd", "bin", "dense_gpu")
CASES = golden_cases()


def _run_driver(tmp_path, c, steps, env=None, **extra):
    import pyoracle as po
    nml = dict(benchmark=c["benchmark"], ni=c["ni"], nj=c["nj"], splitting=c["splitting"], recon=c["recon"],
               cfl=c["cfl"], stages=c.get("stages", 3), if_const_dt=c.get("if_const_dt", False),
               dt=c.get("dt", 0.01), if_bodyforce=c.get("if_bodyforce", False), gamma=c.get("gamma", 1.4))
    nml.update(extra)
    po.write_input_in(str(tmp_path / "input.in"), total_step=steps, save_step=steps, **nml)
    out = subprocess.run([DRIVER], cwd=tmp_path, capture_output=True, text=True, timeout=120,
                         env=None if env is None else {**os.environ, **env})
    assert out.returncode == 0, out.stdout[-2000:] + out.stderr[-2000:]
    dts = [float(m.group(1).replace("D", "E")) for m in re.finditer(r"dt=\s*(\S+)", out.stdout)]
    return po.read_dat2d(str(tmp_path / f"2d_{steps}.dat2d")), np.array(dts), out.stdout


@pytest.mark.parametrize("name", ["dmr_104", "ivc_101_253", "rt_105", "sv_104", "dmr_102_253_constdt",
                                  "ivc_101_255_s2", "sedov_104", "rm_105", "ssl_105_30"])
def test_driver_matches_reference(tmp_path, name):
    if not os.path.exists(DRIVER):
        pytest.fail("Fortran driver not built (python 2dense_amd/build.py)")
    c = CASES[name]
    g = golden(name)
    print(f"driver case {name}", flush=True)
    d, dts, _ = _run_driver(tmp_path, c, c["steps"])
    assert_parity(d["cons"][:, 4:-4, 4:-4], g["cons"], d["dx"], d["dy"], name)
    assert_dts(dts, g["dts"], name)
    assert d["step"] == c["steps"] and d["time"] == pytest.approx(float(g["time"]), rel=1e-12, abs=0)
    if os.path.exists(os.path.join(ROOT, "tests", "golden", f"dump_{name}.npz")):
        # the whole array: ghost layers of the last boundary pass, zero corners
        assert_parity(d["cons"], golden(f"dump_{name}")["cons"], d["dx"], d["dy"], f"{name} dump")
    # the whole file against the reference's dump of the same input.in: the
    # same record layout, header records, cell centres and metrics bit for bit
    import pyoracle as po
    if os.path.exists(po.REF_BIN):
        ref_dir = tmp_path / "ref"
        ref_dir.mkdir()
        (ref_dir / "input.in").write_text((tmp_path / "input.in").read_text())
        subprocess.run([po.REF_BIN], cwd=ref_dir, capture_output=True, check=True, timeout=120)
        f = f"2d_{c['steps']}.dat2d"
        mine, theirs = (tmp_path / f).read_bytes(), (ref_dir / f).read_bytes()
        assert len(mine) == len(theirs)
        r = po.read_dat2d(str(ref_dir / f))
        assert (d["step"], d["ni"], d["nj"], d["dx"], d["dy"]) == (r["step"], r["ni"], r["nj"], r["dx"], r["dy"])
        assert np.array_equal(d["x"], r["x"]) and np.array_equal(d["y"], r["y"])
        assert_parity(d["cons"], r["cons"], d["dx"], d["dy"], f"{name} file")


def test_driver_restart_roundtrip(tmp_path):
    """START_STEP restart from a dump continues bit for bit: it ends where an
    uninterrupted run of the driver ends (the reference's loader is
    unreachable; the driver makes it reachable), and at the reference's state."""
    c = CASES["dmr_104_120x30_20"]
    g = golden("dmr_104_120x30_20")
    (tmp_path / "straight").mkdir()
    straight, _, _ = _run_driver(tmp_path / "straight", c, 20)
    _run_driver(tmp_path, c, 10)  # writes 2d_10.dat2d
    import pyoracle as po
    po.write_input_in(str(tmp_path / "input.in"), benchmark=5, ni=c["ni"], nj=c["nj"], splitting=104, recon=255,
                      cfl=c["cfl"], total_step=20, save_step=20)
    txt = (tmp_path / "input.in").read_text().replace("TOTAL_STEP=20,", "TOTAL_STEP=20, START_STEP=10,")
    (tmp_path / "input.in").write_text(txt)
    out = subprocess.run([DRIVER], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout[-2000:]
    d = po.read_dat2d(str(tmp_path / "2d_20.dat2d"))
    assert np.array_equal(d["cons"][:, 4:-4, 4:-4], straight["cons"][:, 4:-4, 4:-4])
    assert d["time"] == straight["time"]
    assert_parity(d["cons"][:, 4:-4, 4:-4], g["cons"], d["dx"], d["dy"], "restart")


@pytest.mark.parametrize("name", ["dmr_104", "ivc_101_253", "ssl_105_30"])
def test_driver_autotune_keeps_the_bits(tmp_path, name):
    """DENSE_AUTOTUNE=1 makes the driver time the tile shapes before its run
    (dense_autotune restores state and clock); the outputs stay the
    same bit for bit (every tile shape gives the same bits), and the
    reference's to the tolerances."""
    if not os.path.exists(DRIVER):
        pytest.fail("Fortran driver not built (python 2dense_amd/build.py)")
    c = CASES[name]
    steps = c["steps"]
    print(f"driver autotune case {name}", flush=True)
    (tmp_path / "on").mkdir()
    (tmp_path / "off").mkdir()
    on, dts_on, out_on = _run_driver(tmp_path / "on", c, steps, env={"DENSE_AUTOTUNE": "1"})
    off, dts_off, out_off = _run_driver(tmp_path / "off", c, steps, env={"DENSE_AUTOTUNE": "0"})
    assert "Tile shape chosen by timing" in out_on and "Tile shape chosen by timing" not in out_off
    assert np.array_equal(dts_on, dts_off)
    assert np.array_equal(on["cons"], off["cons"])
    g = golden(name)
    assert_dts(dts_on, g["dts"], name)
    assert_parity(on["cons"][:, 4:-4, 4:-4], g["cons"], on["dx"], on["dy"], name)


@pytest.mark.parametrize("name", ["dmr_105", "rt_105", "sv_105", "ssl_105_30"])
def test_driver_theta_file_matches_reference(tmp_path, name):
    """With IF_SAVE_THETA the driver writes 2d_theta_<step>.dat beside its
    Tecplot file, as the reference does (output_test_general.f90:10-12,
    :61-119): AdaWENO's indicator from the last stage's flux evaluation, the
    larger of each cell's two faces per sweep and sign.  The file equals the
    reference binary's: the same header lines and vertex coordinates, and the
    indicator values to 1e-8.  theta = 1/(1 + (sum alpha - 1)^2) is built from
    second differences of the split flux; on nearly smooth data they cancel to
    1e-8 of the flux and amplify a last-bit difference of the flux (the GPU
    contracts multiply-adds) by as much, so theta agrees to ~1e-9 absolute."""
    import pyoracle as po
    if not os.path.exists(DRIVER):
        pytest.fail("Fortran driver not built (python 2dense_amd/build.py)")
    if not os.path.exists(po.REF_BIN):
        pytest.fail("reference binary not built (oracle/build_ref.sh)")
    c = CASES[name]
    steps = c["steps"]
    print(f"theta file case {name}", flush=True)
    nml = dict(benchmark=c["benchmark"], ni=c["ni"], nj=c["nj"], splitting=105, recon=c["recon"], cfl=c["cfl"],
               gamma=c.get("gamma", 1.4), if_bodyforce=c.get("if_bodyforce", False), total_step=steps,
               save_step=steps, if_save_tec=True, if_save_theta=True)
    files = {}
    for who, exe in (("ref", po.REF_BIN), ("drv", DRIVER)):
        d = tmp_path / who
        d.mkdir()
        po.write_input_in(str(d / "input.in"), **nml)
        out = subprocess.run([exe], cwd=d, capture_output=True, text=True, timeout=120)
        assert out.returncode == 0, out.stdout[-2000:] + out.stderr[-2000:]
        files[who] = (d / f"2d_theta_{steps}.dat").read_text().splitlines()
    ref, got = files["ref"], files["drv"]
    assert len(got) == len(ref) == 3 + 2 * (c["ni"] + 1) * (c["nj"] + 1) + 4 * c["ni"] * c["nj"]
    assert got[1:3] == ref[1:3]  # the title line carries the flow time
    # vertex coordinates, then the four indicator fields
    a = np.array([float(x.replace("D", "E")) for x in got[3:]])
    b = np.array([float(x.replace("D", "E")) for x in ref[3:]])
    nv = 2 * (c["ni"] + 1) * (c["nj"] + 1)
    assert np.array_equal(a[:nv], b[:nv])
    np.testing.assert_allclose(a[nv:], b[nv:], rtol=0, atol=1e-8)
    th = a[nv:]
    assert np.all((th > 0) & (th <= 1))


@pytest.mark.parametrize("name", ["dmr_104", "ivc_101_253", "rt_105", "sedov_104"])
def test_driver_tecplot_file_matches_reference(tmp_path, name):
    """With IF_SAVE_TEC the driver's 2d_<step>.dat equals the reference's
    (output_test_general.f90:16-58): the same header lines, including the
    variable list that names 7 fields for 8 data blocks, and the same vertex
    coordinates, and the primitive fields to 1e-11 relative to each field's
    largest value."""
    import pyoracle as po
    if not os.path.exists(DRIVER) or not os.path.exists(po.REF_BIN):
        pytest.fail("driver or reference binary not built")
    c = CASES[name]
    steps = c["steps"]
    print(f"tecplot file case {name}", flush=True)
    nml = dict(benchmark=c["benchmark"], ni=c["ni"], nj=c["nj"], splitting=c["splitting"], recon=c["recon"],
               cfl=c["cfl"], gamma=c.get("gamma", 1.4), if_bodyforce=c.get("if_bodyforce", False),
               stages=c.get("stages", 3), if_const_dt=c.get("if_const_dt", False), dt=c.get("dt", 0.01),
               total_step=steps, save_step=steps, if_save_tec=True)
    files = {}
    for who, exe in (("ref", po.REF_BIN), ("drv", DRIVER)):
        d = tmp_path / who
        d.mkdir()
        po.write_input_in(str(d / "input.in"), **nml)
        out = subprocess.run([exe], cwd=d, capture_output=True, text=True, timeout=120)
        assert out.returncode == 0, out.stdout[-2000:] + out.stderr[-2000:]
        files[who] = (d / f"2d_{steps}.dat").read_text().splitlines()
    ref, got = files["ref"], files["drv"]
    assert len(got) == len(ref) == 3 + 2 * (c["ni"] + 1) * (c["nj"] + 1) + 6 * c["ni"] * c["nj"]
    assert got[1:3] == ref[1:3]  # the title line carries the flow time
    a = np.array([float(x.replace("D", "E")) for x in got[3:]])
    b = np.array([float(x.replace("D", "E")) for x in ref[3:]])
    nv = 2 * (c["ni"] + 1) * (c["nj"] + 1)
    assert np.array_equal(a[:nv], b[:nv])
    n = c["ni"] * c["nj"]
    for k in range(6):
        fa, fb = a[nv + k * n: nv + (k + 1) * n], b[nv + k * n: nv + (k + 1) * n]
        assert np.max(np.abs(fa - fb)) <= 1e-11 * np.max(np.abs(fb)), f"field {k}"


@pytest.mark.parametrize("if_save_tec", [True, False])
def test_driver_save_cadence_matches_reference(tmp_path, if_save_tec):
    """Saves at MOD(step, SAVE_STEP) = 0 or once the save clock reaches DT_SAVE
    (general_main.f90:38-53), the clock reset only when the Tecplot file is
    written (so without IF_SAVE_TEC every step after the first crossing saves):
    the driver writes its dumps at the same steps as the reference, with the
    same interior states and times (to the tolerances)."""
    import pyoracle as po
    if not os.path.exists(DRIVER) or not os.path.exists(po.REF_BIN):
        pytest.fail("driver or reference binary not built")
    c = CASES["dmr_104"]
    g = golden("dmr_104")
    steps = c["steps"]
    nml = dict(benchmark=c["benchmark"], ni=c["ni"], nj=c["nj"], splitting=c["splitting"], recon=c["recon"],
               cfl=c["cfl"], total_step=steps, save_step=7, dt_save=float(g["time"]) / 4.5,
               if_save_tec=if_save_tec)
    dumps = {}
    for who, exe in (("ref", po.REF_BIN), ("drv", DRIVER)):
        d = tmp_path / who
        d.mkdir()
        po.write_input_in(str(d / "input.in"), **nml)
        out = subprocess.run([exe], cwd=d, capture_output=True, text=True, timeout=120)
        assert out.returncode == 0, out.stdout[-2000:] + out.stderr[-2000:]
        dumps[who] = {int(f[3:-6]): po.read_dat2d(str(d / f)) for f in os.listdir(d)
                      if f.startswith("2d_") and f.endswith(".dat2d")}
    print(f"save steps: reference {sorted(dumps['ref'])}, driver {sorted(dumps['drv'])}", flush=True)
    assert sorted(dumps["drv"]) == sorted(dumps["ref"]) and len(dumps["ref"]) >= 2
    for k in dumps["ref"]:
        a, b = dumps["drv"][k], dumps["ref"][k]
        assert a["step"] == b["step"] and a["time"] == pytest.approx(b["time"], rel=1e-12, abs=0)
        assert_parity(a["cons"][:, 4:-4, 4:-4], b["cons"][:, 4:-4, 4:-4], a["dx"], a["dy"], f"save {k}")
