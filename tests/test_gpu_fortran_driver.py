"""The Fortran driver (iso_c_binding over the C ABI) end to end.

The driver reads the reference's input.in, builds the canned problem, steps
it on the GPU and writes 2d_<step>.dat2d in the reference's unformatted
layout; its interior state and printed dt sequence must equal the
reference's own outputs bit for bit.
"""
import os
import re
import subprocess

import numpy as np
import pytest

from conftest import ROOT, golden, golden_cases

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
    got = d["cons"][:, 4:-4, 4:-4]
    assert np.array_equal(got, g["cons"]), f"max |diff| {np.max(np.abs(got - g['cons']))}"
    assert np.array_equal(dts, g["dts"])
    assert d["step"] == c["steps"] and d["time"] == float(g["time"])
    if os.path.exists(os.path.join(ROOT, "tests", "golden", f"dump_{name}.npz")):
        # the whole array: ghost layers of the last boundary pass, zero corners
        assert np.array_equal(d["cons"], golden(f"dump_{name}")["cons"])
    # the whole file, byte for byte, against the reference's dump of the same input.in
    import pyoracle as po
    if os.path.exists(po.REF_BIN):
        ref_dir = tmp_path / "ref"
        ref_dir.mkdir()
        (ref_dir / "input.in").write_text((tmp_path / "input.in").read_text())
        subprocess.run([po.REF_BIN], cwd=ref_dir, capture_output=True, check=True, timeout=120)
        f = f"2d_{c['steps']}.dat2d"
        assert (tmp_path / f).read_bytes() == (ref_dir / f).read_bytes()


def test_driver_restart_roundtrip(tmp_path):
    """START_STEP restart from a dump continues bit for bit (the reference's
    loader is unreachable; the driver makes it reachable)."""
    c = CASES["dmr_104_120x30_20"]
    g = golden("dmr_104_120x30_20")
    _run_driver(tmp_path, c, 10)  # writes 2d_10.dat2d
    import pyoracle as po
    po.write_input_in(str(tmp_path / "input.in"), benchmark=5, ni=c["ni"], nj=c["nj"], splitting=104, recon=255,
                      cfl=c["cfl"], total_step=20, save_step=20)
    txt = (tmp_path / "input.in").read_text().replace("TOTAL_STEP=20,", "TOTAL_STEP=20, START_STEP=10,")
    (tmp_path / "input.in").write_text(txt)
    out = subprocess.run([DRIVER], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout[-2000:]
    d = po.read_dat2d(str(tmp_path / "2d_20.dat2d"))
    assert np.array_equal(d["cons"][:, 4:-4, 4:-4], g["cons"])


@pytest.mark.parametrize("name", ["dmr_104", "ivc_101_253", "ssl_105_30"])
def test_driver_autotune_keeps_the_bits(tmp_path, name):
    """DENSE_AUTOTUNE=1 makes the driver time the tile shapes before its run
    (dense_autotune restores state and clock); the outputs stay the
    reference's, bit for bit, and the printed dt sequence is unchanged."""
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
    g = golden(name)
    assert np.array_equal(dts_on, g["dts"])
    for d in (on, off):
        assert np.array_equal(d["cons"][:, 4:-4, 4:-4], g["cons"])


@pytest.mark.parametrize("name", ["dmr_105", "rt_105", "sv_105", "ssl_105_30"])
def test_driver_theta_file_matches_reference(tmp_path, name):
    """With IF_SAVE_THETA the driver writes 2d_theta_<step>.dat beside its
    Tecplot file, as the reference does (output_test_general.f90:10-12,
    :61-119): AdaWENO's indicator from the last stage's flux evaluation, the
    larger of each cell's two faces per sweep and sign.  The file equals the
    reference binary's, text for text."""
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
    assert got[:3] == ref[:3]
    # vertex coordinates, then the four indicator fields: the same values
    a = np.array([float(x.replace("D", "E")) for x in got[3:]])
    b = np.array([float(x.replace("D", "E")) for x in ref[3:]])
    assert np.array_equal(a, b), f"max |diff| {np.max(np.abs(a - b))}"
    th = a[2 * (c["ni"] + 1) * (c["nj"] + 1):]
    assert np.all((th > 0) & (th <= 1))


@pytest.mark.parametrize("name", ["dmr_104", "ivc_101_253", "rt_105", "sedov_104"])
def test_driver_tecplot_file_matches_reference(tmp_path, name):
    """With IF_SAVE_TEC the driver's 2d_<step>.dat equals the reference's
    (output_test_general.f90:16-58): the same header lines, including the
    variable list that names 7 fields for 8 data blocks, and the same vertex
    coordinates and primitive fields, value for value."""
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
    assert got[:3] == ref[:3]
    a = np.array([float(x.replace("D", "E")) for x in got[3:]])
    b = np.array([float(x.replace("D", "E")) for x in ref[3:]])
    assert np.array_equal(a, b), f"max |diff| {np.max(np.abs(a - b))}"


@pytest.mark.parametrize("if_save_tec", [True, False])
def test_driver_save_cadence_matches_reference(tmp_path, if_save_tec):
    """Saves at MOD(step, SAVE_STEP) = 0 or once the save clock reaches DT_SAVE
    (general_main.f90:38-53), the clock reset only when the Tecplot file is
    written (so without IF_SAVE_TEC every step after the first crossing saves):
    the driver writes its dumps at the same steps as the reference, with the
    same interior states and times."""
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
        assert (a["step"], a["time"]) == (b["step"], b["time"])
        assert np.array_equal(a["cons"][:, 4:-4, 4:-4], b["cons"][:, 4:-4, 4:-4])
