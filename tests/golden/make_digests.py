"""Write digests of the reference's own runs at the workload sizes.

The golden fixtures (make_golden.py) hold whole states of small grids.  The
configured workloads are too large for that, and some run thousands of steps
to their end times (the reference's loop, general_main.f90:23-37, runs while
current_time < total_time with no clipping of the last step).  For those this
script runs the reference solver (compiled from /root/reference/source into
oracle/_ref/ by oracle/build_ref.sh; the race-fixed OpenMP build for the
characteristic schemes 104/105, bitwise equal to the one-thread reference)
and stores per case: the step count, the final time, every dt it printed,
the SHA-256 of the final interior state and a block-averaged density field
(tests/golden/digest.py).  Run it in the build container, where the
reference sources exist; the GPU tests compare against the stored files.

    python tests/golden/make_digests.py [--threads N] [case names]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "oracle"))
sys.path.insert(0, HERE)
import pyoracle as po  # noqa: E402
from digest import coarse_density, state_digest  # noqa: E402

RT = {"if_bodyforce": True, "gamma": 1.6666666666666667}
# name: (benchmark, ni, nj, splitting, recon, cfl, steps, until_time, extra namelist)
CASES = {
    # double Mach reflection as configured, to t = 0.2
    "dmr_104_960x240_t0.2": (5, 960, 240, 104, 255, 0.6, 10 ** 6, 0.2, {}),
    "dmr_104_1920x480_t0.2": (5, 1920, 480, 104, 255, 0.6, 10 ** 6, 0.2, {}),
    # the reference's benchmark 1 (its own vortex, extrapolated x sides) at 200^2, 100 steps
    "ivc_101_253_200x200_100": (1, 200, 200, 101, 253, 0.5, 100, None, {}),
    # Rayleigh-Taylor, AdaWENO with gravity: 1024x4096 for 20 steps, 256x1024 to
    # t = 0.975 and to t = 1.95
    "rt_105_1024x4096_20": (3, 1024, 4096, 105, 255, 0.5, 20, None, RT),
    "rt_105_256x1024_t0.975": (3, 256, 1024, 105, 255, 0.5, 10 ** 6, 0.975, RT),
    "rt_105_256x1024_t1.95": (3, 256, 1024, 105, 255, 0.5, 10 ** 6, 1.95, RT),
}


def run_case(name, threads):
    bm, ni, nj, sp, rc, cfl, steps, until, extra = CASES[name]
    nml = dict(benchmark=bm, ni=ni, nj=nj, splitting=sp, recon=rc, cfl=cfl, threads=threads)
    nml.update(extra)
    binary = po.REF_BIN_RACEFIX if sp in (104, 105) else po.REF_BIN
    t0 = time.time()
    d = po.run_reference(steps=steps, binary=binary, until_time=until, **nml)
    wall = time.time() - t0
    interior = d["cons"][:, po.BF:-po.BF, po.BF:-po.BF]
    dts = np.asarray(d["dts"], dtype=np.float64)
    if len(dts) != d["step"]:
        raise RuntimeError(f"{name}: {len(dts)} dt lines for {d['step']} steps")
    np.savez_compressed(os.path.join(HERE, f"digest_{name}.npz"), dts=dts, coarse=coarse_density(interior),
                        time=np.float64(d["time"]), step=np.int64(d["step"]))
    entry = dict(benchmark=bm, ni=ni, nj=nj, splitting=sp, recon=rc, cfl=cfl, max_steps=steps,
                 until_time=until, step=int(d["step"]), time=float(d["time"]), **extra,
                 **state_digest(interior), reference_threads=threads, reference_wall_s=round(wall, 1))
    print(f"{name}: {d['step']} steps, t = {d['time']!r}, {wall:.0f} s", flush=True)
    return entry


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--threads", type=int, default=os.cpu_count() or 1)
    ap.add_argument("names", nargs="*")
    args = ap.parse_args()
    path = os.path.join(HERE, "digests.json")
    index = {}
    if os.path.exists(path):
        with open(path) as f:
            index = json.load(f)
    for name in args.names or list(CASES):
        index[name] = run_case(name, args.threads)
        with open(path, "w") as f:  # after every case: a long run keeps what finished
            json.dump(index, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
