"""Write the golden fixtures: outputs of the reference itself.

Runs the reference solver (compiled from /root/reference/source into
oracle/_ref/ by oracle/build_ref.sh) through its own input.in and .dat2d
formats on small grids, for every splitting x reconstruction switch the hot
path implements, and stores the interior conserved state after N steps plus
the per-step dt it printed.  These files are data (inputs = the switches in
`CASES`, outputs = the reference's numbers); run this script in the build
container, where the reference sources exist.

    python tests/golden/make_golden.py [case names]   # all cases, or only those named
    python tests/golden/make_golden.py --dumps [names] # whole dumps of DUMP_CASES
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "oracle"))
import pyoracle as po  # noqa: E402

# (name, benchmark, ni, nj, splitting, recon, cfl, steps, extra namelist)
CASES = []
for sp in (101, 102, 103):
    for rc in (251, 253, 255):
        CASES.append((f"ivc_{sp}_{rc}", 1, 32, 32, sp, rc, 0.5, 5, {}))
CASES += [
    ("ivc_104", 1, 32, 32, 104, 255, 0.5, 5, {}),
    ("ivc_105", 1, 32, 32, 105, 255, 0.5, 5, {}),
    ("ivc_101_253_s1", 1, 32, 32, 101, 253, 0.5, 5, {"stages": 1}),
    ("ivc_101_255_s2", 1, 32, 32, 101, 255, 0.5, 5, {"stages": 2}),
    ("ivc_101_253_200", 1, 200, 200, 101, 253, 0.5, 3, {}),
    ("dmr_101_253", 5, 64, 16, 101, 253, 0.6, 10, {}),
    ("dmr_102_255", 5, 64, 16, 102, 255, 0.6, 10, {}),
    ("dmr_103_251", 5, 64, 16, 103, 251, 0.6, 10, {}),
    ("dmr_104", 5, 64, 16, 104, 255, 0.6, 10, {}),
    ("dmr_105", 5, 64, 16, 105, 255, 0.6, 10, {}),
    ("dmr_104_120x30_20", 5, 120, 30, 104, 255, 0.6, 20, {}),
    ("dmr_102_253_constdt", 5, 64, 16, 102, 253, 0.6, 10, {"if_const_dt": True, "dt": 1.0e-4}),
    ("rt_104", 3, 16, 64, 104, 255, 0.5, 10, {"if_bodyforce": True, "gamma": 1.6666666666666667}),
    ("rt_105", 3, 16, 64, 105, 255, 0.5, 10, {"if_bodyforce": True, "gamma": 1.6666666666666667}),
    ("rt_101_255", 3, 16, 64, 101, 255, 0.5, 10, {"if_bodyforce": True, "gamma": 1.6666666666666667}),
    ("sv_101_255", 9, 64, 32, 101, 255, 0.5, 10, {}),
    ("sv_104", 9, 64, 32, 104, 255, 0.5, 10, {}),
    ("sv_105", 9, 64, 32, 105, 255, 0.5, 10, {}),
    # the characteristic schemes ignore the reconstruction switch (always WENO-Z,
    # scheme_weno5th_CH.f90:73, scheme_weno5th_SC.f90:24-28)
    ("dmr_104_recon251", 5, 64, 16, 104, 251, 0.6, 10, {}),
    ("sv_105_recon253", 9, 64, 32, 105, 253, 0.5, 10, {}),
    ("sedov_101_255", 2, 48, 48, 101, 255, 0.5, 10, {}),
    ("sedov_104", 2, 48, 48, 104, 255, 0.5, 10, {}),
    ("sedov_103_253", 2, 40, 40, 103, 253, 0.5, 10, {}),
    ("rm_101_253", 4, 64, 16, 101, 253, 0.5, 10, {}),
    ("rm_104", 4, 64, 16, 104, 255, 0.5, 10, {}),
    ("rm_105", 4, 64, 16, 105, 255, 0.5, 10, {}),
    ("ssl_101_255", 6, 64, 32, 101, 255, 0.5, 10, {}),
    ("ssl_104", 6, 64, 32, 104, 255, 0.5, 10, {}),
    ("ssl_105_30", 6, 80, 16, 105, 255, 0.5, 30, {}),
    ("ssl_102_253", 6, 64, 32, 102, 253, 0.5, 10, {}),
]


# Cases whose whole dump (ghost layers and corners included) is also stored:
# the reference's cons array at a save holds the ghost values of its last
# boundary pass (the last stage's input at t^n), general_main.f90:26-31,43-44.
DUMP_CASES = ["ivc_101_253", "ivc_101_255_s2", "dmr_104", "rt_105", "sv_101_255", "ssl_101_255",
              "sedov_101_255", "rm_104"]


def write_dumps(only=None):
    for name, bm, ni, nj, sp, rc, cfl, steps, extra in CASES:
        if name not in DUMP_CASES or (only and name not in only):
            continue
        nml = dict(benchmark=bm, ni=ni, nj=nj, splitting=sp, recon=rc, cfl=cfl)
        nml.update(extra)
        d = po.run_reference(steps=steps, **nml)
        np.savez_compressed(os.path.join(HERE, f"dump_{name}.npz"), cons=d["cons"], step=np.int64(d["step"]),
                            time=np.float64(d["time"]))
        print("dump", name, d["cons"].shape)


def main(only=None):
    """only: names of the cases to (re)write; the index lists every case."""
    index = {}
    for name, bm, ni, nj, sp, rc, cfl, steps, extra in CASES:
        nml = dict(benchmark=bm, ni=ni, nj=nj, splitting=sp, recon=rc, cfl=cfl)
        nml.update(extra)
        index[name] = dict(nml, steps=steps)
        if only and name not in only:
            continue
        d = po.run_reference(steps=steps, **nml)
        cons = d["cons"][:, po.BF:-po.BF, po.BF:-po.BF]
        np.savez_compressed(os.path.join(HERE, f"{name}.npz"), cons=cons, dts=d["dts"],
                            time=np.float64(d["time"]), step=np.int64(d["step"]))
        print(name, cons.shape, "t =", d["time"])
    with open(os.path.join(HERE, "cases.json"), "w") as f:
        json.dump(index, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    if sys.argv[1:2] == ["--dumps"]:
        write_dumps(set(sys.argv[2:]) or None)
    else:
        main(set(sys.argv[1:]) or None)
