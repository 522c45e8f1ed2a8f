"""Algorithmic FP64 operations per cell-update, counted by the oracle.

Runs the op-counting build of the C restatement (oracle/_build/liboracle_ops.so,
single thread) for one time step on a given problem and divides the count of
+,-,*,/,sqrt executed by ni*nj (one cell-update = one interior cell advanced
one full 3-stage RK step).  Loop-invariant parameter expressions are not
counted; abs/neg/max/compare count zero.  The count includes the reference's
own boundary overhead (fluxes on the ghost columns, ni+1 faces per row), so
it is quoted at the workload's grid size.

    python oracle/count_ops.py 5 1920 480 104 255
"""
import json
import sys

import pyoracle as po


def count(benchmark, ni, nj, splitting, recon, steps=1, warm=0):
    p = po.problem_params(benchmark, ni, nj, splitting=splitting, recon=recon,
                          cfl=0.6 if benchmark == 5 else 0.5, ops=True)
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
    print(json.dumps({"benchmark": bm, "ni": ni, "nj": nj, "splitting": sp, "recon": rc,
                      "flops_per_cell_update": count(bm, ni, nj, sp, rc)}))
