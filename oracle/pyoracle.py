"""Python access to the oracle -- TEST INFRASTRUCTURE ONLY.

Only tests/, ``__graft_entry__.smoke()`` and bench.py's cpu_baseline leg
import this module.  It offers:

* ``Oracle``: ctypes binding of the C restatement (``oracle/_build/liboracle.so``),
  whose functions restate the reference's hot path routine by routine;
* ``write_input_in`` / ``read_dat2d`` / ``run_reference``: drive the reference
  itself, compiled from its sources into ``oracle/_ref/`` by ``build_ref.sh``,
  through its own ``input.in`` NAMELIST file (flow_mod_general.f90:297-334) and
  its unformatted ``2d_<step>.dat2d`` dump (read_WriteAndLoad.f90:1-28).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle.so")
LIB_OPS_PATH = os.path.join(HERE, "_build", "liboracle_ops.so")
REF_BIN = os.path.join(HERE, "_ref", "2dense_ref")
REF_BIN_RACEFIX = os.path.join(HERE, "_ref", "2dense_ref_racefix")

BF = 4
BC_PERIODIC, BC_EXTRAP, BC_REFLECT, BC_DIRICHLET, BC_DMR_WALL, BC_DMR_TOP = range(6)


class OraParams(C.Structure):
    _fields_ = [
        ("ni", C.c_int), ("nj", C.c_int),
        ("splitting", C.c_int), ("recon", C.c_int), ("stages", C.c_int),
        ("if_const_dt", C.c_int), ("if_bodyforce", C.c_int),
        ("dx", C.c_double), ("dy", C.c_double),
        ("gamma", C.c_double), ("cfl", C.c_double), ("delta", C.c_double), ("dt_const", C.c_double),
        ("bc", C.c_int * 4),
        ("bc_state", (C.c_double * 4) * 4),
        ("nthreads", C.c_int),
    ]


_libs: dict = {}


def load(ops: bool = False) -> C.CDLL:
    path = LIB_OPS_PATH if ops else LIB_PATH
    if path in _libs:
        return _libs[path]
    if not os.path.exists(path):
        raise FileNotFoundError(f"oracle library not built: {path} (run `make -C oracle`)")
    lib = C.CDLL(path)
    P = C.POINTER(OraParams)
    dp = C.POINTER(C.c_double)
    vp = C.c_void_p
    lib.ora_create.argtypes = [P]; lib.ora_create.restype = vp
    lib.ora_destroy.argtypes = [vp]
    lib.ora_problem_params.argtypes = [C.c_int, C.c_int, C.c_int, P]; lib.ora_problem_params.restype = C.c_int
    lib.ora_init_state.argtypes = [vp, C.c_int]; lib.ora_init_state.restype = C.c_int
    for n in ("ora_set_cons", "ora_get_cons", "ora_get_prim", "ora_set_bodyforce"):
        getattr(lib, n).argtypes = [vp, dp]
    lib.ora_step.argtypes = [vp]
    lib.ora_run.argtypes = [vp, C.c_double, C.c_int]; lib.ora_run.restype = C.c_int
    lib.ora_time.argtypes = [vp]; lib.ora_time.restype = C.c_double
    lib.ora_set_time.argtypes = [vp, C.c_double]
    lib.ora_current_step.argtypes = [vp]; lib.ora_current_step.restype = C.c_int
    lib.ora_dt.argtypes = [vp]; lib.ora_dt.restype = C.c_double
    lib.ora_set_dt.argtypes = [vp, C.c_double]
    lib.ora_max_eig.argtypes = [vp, dp]
    lib.ora_update_params.argtypes = [vp, C.c_double, C.c_double, C.c_double, C.c_int, C.c_double, C.c_int]
    for n in ("ora_weno5z_stencil", "ora_weno5js_stencil", "ora_upwind5_stencil"):
        getattr(lib, n).argtypes = [dp]; getattr(lib, n).restype = C.c_double
    lib.ora_weno5z_omega_theta.argtypes = [dp, dp]
    lib.ora_lr_vec_xi.argtypes = [C.c_double] * 5 + [dp, dp]
    lib.ora_lr_vec_eta.argtypes = [C.c_double] * 5 + [dp, dp]
    lib.ora_op_count.restype = C.c_longlong
    lib.ora_set_mutation.argtypes = [C.c_int]
    lib.ora_lf_split.argtypes = [C.c_double, C.c_double, C.c_double, dp, C.c_int, dp, dp]
    lib.ora_sw_split.argtypes = [C.c_double] * 5 + [C.c_int, dp, dp]
    _libs[path] = lib
    return lib


class planted_mistake:
    """Context manager for the mutation tests: the restatement with one planted
    mistake (g_mut in dense_oracle.c), restored on exit."""

    def __init__(self, m: int):
        self.m = m

    def __enter__(self):
        load().ora_set_mutation(self.m)
        return self

    def __exit__(self, *exc):
        load().ora_set_mutation(0)


def _dptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"] or a.flags["F_CONTIGUOUS"]
    return a.ctypes.data_as(C.POINTER(C.c_double))


def problem_params(benchmark: int, ni: int, nj: int, *, splitting=101, recon=255, cfl=0.5,
                   gamma=1.4, delta=0.5, stages=3, if_const_dt=0, dt_const=0.01,
                   nthreads=1, ops=False) -> OraParams:
    """Parameters for one of the reference's canned problems (init_general.f90:13-37)."""
    lib = load(ops)
    p = OraParams()
    p.gamma = gamma
    rc = lib.ora_problem_params(benchmark, ni, nj, C.byref(p))
    if rc != 0:
        raise ValueError(f"benchmark {benchmark} not restated")
    p.splitting, p.recon, p.stages = splitting, recon, stages
    p.cfl, p.delta = cfl, delta
    p.if_const_dt, p.dt_const = if_const_dt, dt_const
    p.nthreads = nthreads
    return p


class Oracle:
    """One block of the CPU restatement (same layout as the reference)."""

    def __init__(self, params: OraParams, ops: bool = False):
        self.lib = load(ops)
        self.p = params
        self.h = self.lib.ora_create(C.byref(params))
        if not self.h:
            raise MemoryError("ora_create failed")
        self.ni, self.nj = params.ni, params.nj

    def close(self):
        if self.h:
            self.lib.ora_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def shape_cons(self):
        return (4, self.nj + 2 * BF, self.ni + 2 * BF)

    def init_state(self, benchmark: int):
        if self.lib.ora_init_state(self.h, benchmark) != 0:
            raise ValueError(f"benchmark {benchmark} has no restated initial state")

    def set_cons(self, cons: np.ndarray):
        a = np.ascontiguousarray(cons, dtype=np.float64).reshape(self.shape_cons)
        self.lib.ora_set_cons(self.h, _dptr(a))

    def set_bodyforce(self, bf: np.ndarray):
        a = np.ascontiguousarray(bf, dtype=np.float64).reshape(self.shape_cons)
        self.lib.ora_set_bodyforce(self.h, _dptr(a))

    def cons(self) -> np.ndarray:
        """Padded conserved state as (4, nj+8, ni+8): the Fortran cons(ni+8,nj+8,4) memory."""
        a = np.empty(self.shape_cons)
        self.lib.ora_get_cons(self.h, _dptr(a))
        return a

    def prim(self) -> np.ndarray:
        a = np.empty((6, self.nj + 2 * BF, self.ni + 2 * BF))
        self.lib.ora_get_prim(self.h, _dptr(a))
        return a

    def interior(self) -> np.ndarray:
        return self.cons()[:, BF:-BF, BF:-BF].copy()

    def step(self, n: int = 1):
        for _ in range(n):
            self.lib.ora_step(self.h)

    def run(self, total_time: float, total_step: int) -> int:
        return self.lib.ora_run(self.h, total_time, total_step)

    @property
    def time(self) -> float:
        return self.lib.ora_time(self.h)

    @time.setter
    def time(self, t: float):
        self.lib.ora_set_time(self.h, t)

    @property
    def dt(self) -> float:
        return self.lib.ora_dt(self.h)

    @property
    def current_step(self) -> int:
        return self.lib.ora_current_step(self.h)

    def max_eig(self):
        a = np.empty(2)
        self.lib.ora_max_eig(self.h, _dptr(a))
        return float(a[0]), float(a[1])

    def op_count(self) -> int:
        return int(self.lib.ora_op_count())


# --------------------------------------------------------------------------
# The reference binary, driven through its own input.in / .dat2d formats
# --------------------------------------------------------------------------

def write_input_in(path: str, *, benchmark: int, ni: int, nj: int, splitting: int, recon: int,
                   cfl: float, total_step: int, save_step: int, total_time: float = 1.0e30,
                   gamma: float = 1.4, delta: float = 0.5, threads: int = 1, stages: int = 3,
                   if_const_dt: bool = False, dt: float = 0.01, if_bodyforce: bool = False,
                   if_save_binary: bool = True, if_save_tec: bool = False, dt_save: float = 1.0e30,
                   if_save_theta: bool = False):
    """Six NAMELIST groups in the order read_par_namelist reads them (flow_mod_general.f90:313-332)."""
    L = lambda b: "T" if b else "F"  # noqa: E731
    txt = f"""&OMP_PARS
 OMP_THREAD_NUM={threads},
 /
&MODEL_PARS
 NEQ=4, IF_SOURCE=F, IF_VISCOUS=F, IF_BODYFORCE={L(if_bodyforce)}, MACH_INF=1.0, REYNOLDS=500.0,
 BENCHMARK_INDEX={benchmark},
 /
&GRID_PARS
 IF_CURVE_GRID=F, IF_READ_FROM_FILE=F, SINGLE_NI={ni}, SINGLE_NJ={nj},
 /
&TIME_STEP_PARS
 STAGES={stages}, TOTAL_TIME={total_time!r}, DT_SAVE={dt_save!r}, DT={dt!r}, CFL={cfl!r},
 TOTAL_STEP={total_step}, SAVE_STEP={save_step},
 IF_CONST_DT={L(if_const_dt)}, IF_SAVE_BINARY={L(if_save_binary)}, IF_SAVE_TEC={L(if_save_tec)},
 /
&SCHEME_PARS
 SCHEME_SPLITTING_INDEX={splitting}, SCHEME_INV_INDEX={recon}, SCHEME_VIS_INDEX=301,
 IF_SAVE_THETA={L(if_save_theta)}, DELTA={delta!r},
 /
&GAS_PARS
 GAMMA={gamma!r},
 /
"""
    txt = txt.replace("e+", "d+").replace("e-", "d-") if False else txt
    with open(path, "w") as f:
        f.write(txt)


def _records(buf: bytes):
    off = 0
    while off < len(buf):
        n = int(np.frombuffer(buf, dtype="<i4", count=1, offset=off)[0])
        yield buf[off + 4: off + 4 + n]
        off += 8 + n


def read_dat2d(path: str) -> dict:
    """Parse the unformatted-sequential dump written by write_date_binary
    (read_WriteAndLoad.f90:1-28): (step,time), num_block, index, (ni,nj),
    (bf_i,bf_j), cons incl. ghosts, x, y, (dx,dy)."""
    with open(path, "rb") as f:
        recs = list(_records(f.read()))
    step = int(np.frombuffer(recs[0], "<i4", 1, 0)[0])
    time = float(np.frombuffer(recs[0], "<f8", 1, 4)[0])
    ni, nj = (int(v) for v in np.frombuffer(recs[3], "<i4", 2))
    bfi, bfj = (int(v) for v in np.frombuffer(recs[4], "<i4", 2))
    cons = np.frombuffer(recs[5], "<f8").reshape(4, nj + 2 * bfj, ni + 2 * bfi).copy()
    x = np.frombuffer(recs[6], "<f8").reshape(nj, ni).copy()
    y = np.frombuffer(recs[7], "<f8").reshape(nj, ni).copy()
    dx, dy = (float(v) for v in np.frombuffer(recs[8], "<f8", 2))
    return dict(step=step, time=time, ni=ni, nj=nj, cons=cons, x=x, y=y, dx=dx, dy=dy)


def run_reference(*, steps: int, binary: str = REF_BIN, workdir: str | None = None,
                  until_time: float | None = None, **nml) -> dict:
    """Run the reference for `steps` steps and return its .dat2d dump at that step,
    plus its stdout (per-step dt).  With until_time, run while t < until_time
    (at most `steps` steps) instead: DT_SAVE = TOTAL_TIME makes the reference's
    save clock (elasp_time, general_main.f90:38) reach it at the same step as
    the flow clock, so it dumps exactly the last step; that dump is returned."""
    if not os.path.exists(binary):
        raise FileNotFoundError(f"reference binary not built: {binary} (run oracle/build_ref.sh)")
    own = workdir is None
    wd = tempfile.mkdtemp(prefix="dense_ref_") if own else workdir
    try:
        if until_time is None:
            write_input_in(os.path.join(wd, "input.in"), total_step=steps, save_step=steps, **nml)
        else:
            write_input_in(os.path.join(wd, "input.in"), total_step=steps, save_step=steps + 1,
                           total_time=until_time, dt_save=until_time, **nml)
        env = dict(os.environ)
        out = subprocess.run([binary], cwd=wd, capture_output=True, text=True, env=env, check=True)
        if until_time is not None:
            dumps = [f for f in os.listdir(wd) if f.startswith("2d_") and f.endswith(".dat2d")]
            last = max(int(f[3:-6]) for f in dumps)
            d = read_dat2d(os.path.join(wd, f"2d_{last}.dat2d"))
        else:
            d = read_dat2d(os.path.join(wd, f"2d_{steps}.dat2d"))
        dts = []
        for line in out.stdout.splitlines():
            if "Flow time=" in line and "dt=" in line:
                dts.append(float(line.split("dt=")[1].replace("D", "E")))
        d["dts"] = np.array(dts)
        d["stdout"] = out.stdout
        return d
    finally:
        if own:
            import shutil
            shutil.rmtree(wd, ignore_errors=True)
