"""Algorithmic FP64 operations per cell-update, counted by the oracle.

Runs the op-counting build of the C restatement (oracle/_build/liboracle_ops.so,
single thread) for one time step on a given problem and divides the count of
+,-,*,/,sqrt executed by ni*nj (one cell-update = one interior cell advanced
one full 3-stage RK step).  Loop-invariant parameter expressions are not
counted; abs/neg/max/compare count zero.  The count includes the reference's
own boundary overhead (fluxes on the ghost columns, ni+1 faces per row), so
it is quoted at the workload's grid size.

    python oracle/count_ops.py 5 1920 480 104 255 [delta]

For AdaWENO (105) the count depends on how many (face, sign) pairs take the
characteristic branch (theta <= delta): delta = 0 sends every pair down the
component-wise branch and delta = 1 every pair down the characteristic one
(theta <= 1 always), and the reference's cost is linear in the fraction of
characteristic pairs between those two counts.
"""
import json
import sys

import pyoracle as po


def count(benchmark, ni, nj, splitting, recon, steps=1, warm=0, delta=0.5):
    p = po.problem_params(benchmark, ni, nj, splitting=splitting, recon=recon,
                          cfl=0.6 if benchmark == 5 else 0.5, delta=delta, ops=True)
    o = po.Oracle(p, ops=True)
    o.init_state(benchmark)
    o.step(warm)
    total = 0
    for _ in range(steps):
        o.step()
        total += o.op_count()
    return total / (ni * nj * steps)


if __name__ == "__main__":
    bm, ni, nj, sp, rc = (int(v) for v in sys.argv[1:6])
    delta = float(sys.argv[6]) if len(sys.argv) > 6 else 0.5
    print(json.dumps({"benchmark": bm, "ni": ni, "nj": nj, "splitting": sp, "recon": rc, "delta": delta,
                      "flops_per_cell_update": count(bm, ni, nj, sp, rc, delta=delta)}))
