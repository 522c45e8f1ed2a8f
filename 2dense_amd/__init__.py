"""MI355X-native explicit update of 2DENSE (FP64), host-side Python binding.

The product is ``lib2dense_amd.so`` (HIP kernels for gfx950 behind the C ABI
declared in ``include/dense_capi.h``).  This module binds that ABI with
ctypes; it allocates nothing on the device itself and has no CPU fallback:
if the library is missing it raises.

The reference keeps the solver state in Fortran module globals and advances
it with four parameterless calls per Runge-Kutta stage
(source/utility/general_main.f90:26-31).  ``Solver`` holds the same state
(``cons`` in the reference layout ``cons(ni+8, nj+8, 4)``, the clock
``current_time``/``current_step``, ``dt``, ``max_eig_i/j``) and ``step()``
performs one full time step on the GPU.

Import as ``importlib.import_module("2dense_amd")`` (the directory name
starts with a digit).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib2dense_amd.so")
BF = 4

# switch codes (flux_splitting.f90:5-17, scheme_invs_general.f90:5-10)
LLF, GLF, SW, CH, SC = 101, 102, 103, 104, 105
UPWIND5, WENOJS, WENOZ = 251, 253, 255
SPLITTINGS = (LLF, GLF, SW, CH, SC)
RECONS = (UPWIND5, WENOJS, WENOZ)

# boundary kinds / sides (include/dense_capi.h)
BC_PERIODIC, BC_EXTRAP, BC_REFLECT, BC_DIRICHLET, BC_DMR_WALL, BC_DMR_TOP = range(6)
LEFT, RIGHT, BOTTOM, TOP = range(4)

DENSE_OK = 0
ERRORS = {-1: "bad argument", -2: "unsupported switch", -3: "HIP error", -4: "non-finite state", -5: "out of memory"}


class DenseError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


class Cfg(C.Structure):
    """dense_cfg of include/dense_capi.h."""
    _fields_ = [
        ("ni", C.c_int), ("nj", C.c_int), ("nghost", C.c_int), ("neq", C.c_int),
        ("dx", C.c_double), ("dy", C.c_double),
        ("gamma", C.c_double), ("cfl", C.c_double), ("delta", C.c_double), ("dt", C.c_double),
        ("splitting", C.c_int), ("recon", C.c_int), ("stages", C.c_int),
        ("if_const_dt", C.c_int), ("if_bodyforce", C.c_int), ("benchmark", C.c_int),
        ("bc", C.c_int * 4),
        ("bc_state", (C.c_double * 4) * 4),
        ("tile", C.c_int),
    ]


_lib = None


def load_library(path: str | None = None) -> C.CDLL:
    """Load lib2dense_amd.so.  Raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    p = path or os.environ.get("DENSE_LIB", LIB_PATH)
    if not os.path.exists(p):
        raise FileNotFoundError(f"{p} is not built; run `python 2dense_amd/build.py`")
    # PyTorch-ROCm ships its own HIP runtime under the same soname
    # (libamdhip64.so.7).  Loaded first, it is the one runtime of the process
    # and the library binds to it; loaded after the library, torch would bring
    # up a second runtime on the same device and find no GPU.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = C.CDLL(p)
    H = C.c_void_p
    dp = C.POINTER(C.c_double)
    P = C.POINTER(Cfg)
    lib.dense_problem_defaults.argtypes = [C.c_int, C.c_int, C.c_int, P]
    lib.dense_create.argtypes = [P, C.POINTER(H)]
    lib.dense_destroy.argtypes = [H]
    lib.dense_set_state.argtypes = [H, dp]
    lib.dense_get_state.argtypes = [H, dp, dp]
    lib.dense_get_dump.argtypes = [H, dp]
    lib.dense_step.argtypes = [H, C.c_double, dp]
    lib.dense_run.argtypes = [H, C.c_int, C.c_double, C.c_int, dp, C.POINTER(C.c_int), dp, C.POINTER(C.c_int)]
    lib.dense_run_saving.argtypes = [H, C.c_int, C.c_double, C.c_int, C.c_double, dp, dp, C.POINTER(C.c_int), dp,
                                     C.POINTER(C.c_int)]
    lib.dense_max_eig.argtypes = [H, dp]
    lib.dense_last_run_gpu_ms.argtypes = [H, dp, C.POINTER(C.c_int)]
    lib.dense_update_params.argtypes = [H, P]
    lib.dense_autotune.argtypes = [H, C.c_int, C.POINTER(C.c_int), dp]
    lib.dense_get_theta.argtypes = [H, dp]
    lib.dense_sync.argtypes = [H]
    lib.dense_probe_fp64.argtypes = [dp]
    lib.dense_probe_hbm.argtypes = [C.c_size_t, dp, dp]
    lib.dense_last_error.argtypes = [H]
    lib.dense_last_error.restype = C.c_char_p
    for name in ("dense_problem_defaults", "dense_create", "dense_destroy", "dense_set_state",
                 "dense_get_state", "dense_get_dump", "dense_step", "dense_run", "dense_run_saving", "dense_max_eig",
                 "dense_last_run_gpu_ms", "dense_update_params", "dense_autotune", "dense_get_theta", "dense_sync",
                 "dense_probe_fp64", "dense_probe_hbm"):
        getattr(lib, name).restype = C.c_int
    _lib = lib
    return lib


def _check(lib, h, rc):
    if rc != DENSE_OK:
        msg = lib.dense_last_error(h)
        raise DenseError(rc, msg.decode() if msg else "")


def _dp(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def make_cfg(ni: int, nj: int, *, benchmark: int = 0, splitting: int = LLF, recon: int = WENOZ,
             cfl: float = 0.5, gamma: float | None = None, delta: float = 0.5, dt: float = 0.01,
             stages: int = 3, if_const_dt: bool = False, if_bodyforce: bool | None = None,
             dx: float | None = None, dy: float | None = None, bc=None, bc_state=None,
             tile: int = 0) -> Cfg:
    """Build a dense_cfg.  For benchmark 1/2/3/4/5/6/9 the grid metrics and boundary
    rules come from the reference's canned problem (dense_problem_defaults;
    benchmark 3 also sets gamma = 5/3 and the gravity body force, as
    init_RT_general.f90:9-10 does); for benchmark 0 give dx, dy, bc (4 kinds)
    and bc_state (4 x (rho,u,v,p)).  gamma=None keeps the problem's value (1.4
    unless the problem sets one).  tile: 0 is the automatic shape (16x15),
    1-3 force 32x7, 8x7, 32x15, 4 forces 16x15 and 5 16x15 with 512 threads
    (x and y faces on separate halves); Solver.autotune picks one by timing;
    results do not depend on it."""
    lib = load_library()
    c = Cfg()
    c.gamma = 1.4
    c.stages = stages
    _check(lib, None, lib.dense_problem_defaults(benchmark, ni, nj, C.byref(c)))
    if gamma is not None:
        c.gamma = gamma
    c.splitting, c.recon = splitting, recon
    c.cfl, c.delta, c.dt = cfl, delta, dt
    c.if_const_dt = int(bool(if_const_dt))
    if if_bodyforce is not None:
        c.if_bodyforce = int(bool(if_bodyforce))
    if dx is not None:
        c.dx = dx
    if dy is not None:
        c.dy = dy
    if bc is not None:
        for s in range(4):
            c.bc[s] = int(bc[s])
    if bc_state is not None:
        for s in range(4):
            for v in range(4):
                c.bc_state[s][v] = float(bc_state[s][v])
    c.tile = tile
    return c


class Solver:
    """One block advanced on the GPU through the C ABI."""

    def __init__(self, cfg: Cfg):
        self.lib = load_library()
        self.cfg = cfg
        h = C.c_void_p()
        _check(self.lib, None, self.lib.dense_create(C.byref(cfg), C.byref(h)))
        self.h = h
        self.ni, self.nj = cfg.ni, cfg.nj
        self.current_time = 0.0
        self.current_step = 0
        self.dt = 0.0

    # --- lifetime
    def close(self):
        if getattr(self, "h", None):
            self.lib.dense_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # --- state
    @property
    def padded_shape(self):
        return (4, self.nj + 2 * BF, self.ni + 2 * BF)

    def set_state(self, cons: np.ndarray):
        """cons: reference layout cons(ni+8, nj+8, 4), i.e. a C array (4, nj+8, ni+8),
        or an interior array (4, nj, ni)."""
        a = np.asarray(cons, dtype=np.float64)
        if a.shape == (4, self.nj, self.ni):
            p = np.zeros(self.padded_shape)
            p[:, BF:-BF, BF:-BF] = a
            a = p
        a = np.ascontiguousarray(a.reshape(self.padded_shape))
        _check(self.lib, self.h, self.lib.dense_set_state(self.h, _dp(a)))

    def get_state(self, with_prim: bool = False):
        cons = np.empty(self.padded_shape)
        prim = np.empty((6,) + self.padded_shape[1:]) if with_prim else None
        _check(self.lib, self.h, self.lib.dense_get_state(self.h, _dp(cons), _dp(prim) if with_prim else None))
        return (cons, prim) if with_prim else cons

    def get_dump(self) -> np.ndarray:
        """cons(ni+8, nj+8, 4) as the reference's array holds it at a .dat2d dump:
        the current interior, ghost layers of the last stage's input at t^n."""
        cons = np.empty(self.padded_shape)
        _check(self.lib, self.h, self.lib.dense_get_dump(self.h, _dp(cons)))
        return cons

    def interior(self) -> np.ndarray:
        return self.get_state()[:, BF:-BF, BF:-BF].copy()

    # --- time stepping
    def step(self, n: int = 1) -> float:
        """n time steps with the reference's clock semantics (general_main.f90:26-37)."""
        dt = C.c_double(0.0)
        for _ in range(n):
            _check(self.lib, self.h, self.lib.dense_step(self.h, self.current_time, C.byref(dt)))
            self.dt = dt.value
            self.current_step += 1
            self.current_time += self.dt
        return self.dt

    def run(self, max_steps: int, total_time: float = float("inf"), total_step: int = 2 ** 31 - 1,
            want_dts: bool = False):
        """Up to max_steps steps while t < total_time and step < total_step, from a
        captured graph with the clock on the device.  Returns steps taken (and dts)."""
        t = C.c_double(self.current_time)
        st = C.c_int(self.current_step)
        taken = C.c_int(0)
        hist = np.zeros(max(1, max_steps)) if want_dts else None
        _check(self.lib, self.h, self.lib.dense_run(self.h, max_steps, total_time, total_step, C.byref(t),
                                                    C.byref(st), _dp(hist) if want_dts else None, C.byref(taken)))
        self.current_time, self.current_step = t.value, st.value
        if want_dts:
            return taken.value, hist[: taken.value].copy()
        return taken.value

    def run_saving(self, max_steps: int, dt_save: float, elasp: float, total_time: float = float("inf"),
                   total_step: int = 2 ** 31 - 1):
        """run() that also stops after the first step at which the save clock
        elasp (advanced by dt each step) reaches dt_save.  Returns (steps taken,
        their dts, the advanced save clock)."""
        t = C.c_double(self.current_time)
        st = C.c_int(self.current_step)
        el = C.c_double(elasp)
        taken = C.c_int(0)
        hist = np.zeros(max(1, max_steps))
        _check(self.lib, self.h, self.lib.dense_run_saving(self.h, max_steps, total_time, total_step, dt_save,
                                                           C.byref(el), C.byref(t), C.byref(st), _dp(hist),
                                                           C.byref(taken)))
        self.current_time, self.current_step = t.value, st.value
        return taken.value, hist[: taken.value].copy(), el.value

    def last_run_gpu_ms(self):
        """(GPU milliseconds of the last run(), RK-stage launches that did work)."""
        ms = C.c_double(0.0)
        n = C.c_int(0)
        _check(self.lib, self.h, self.lib.dense_last_run_gpu_ms(self.h, C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def max_eig(self):
        out = np.zeros(2)
        _check(self.lib, self.h, self.lib.dense_max_eig(self.h, _dp(out)))
        return float(out[0]), float(out[1])

    def update_params(self, **kw):
        for k, v in kw.items():
            setattr(self.cfg, k, v)
        _check(self.lib, self.h, self.lib.dense_update_params(self.h, C.byref(self.cfg)))

    def autotune(self, trial_steps: int = 16):
        """Time each tile shape on trial_steps steps of the current state (which
        is restored) and keep the fastest; returns (cfg.tile value, trial ms).
        Runs replay 16-step graphs, so trials of 16 steps time whole replays."""
        t = C.c_int(0)
        ms = C.c_double(0.0)
        _check(self.lib, self.h, self.lib.dense_autotune(self.h, trial_steps, C.byref(t), C.byref(ms)))
        self.cfg.tile = t.value
        return t.value, ms.value

    def theta(self):
        """AdaWENO's indicator at every face (splitting 105, stages 2-3, after a
        step), as the reference's theta arrays hold it at output time: xp, xn of
        shape (nj, ni+1) and ep, en of shape (ni, nj+1) -- the Fortran arrays
        theta_xi(ni+1, nj) and theta_eta(nj+1, ni) seen in row-major order."""
        ni, nj = self.cfg.ni, self.cfg.nj
        nx, ny = (ni + 1) * nj, (nj + 1) * ni
        buf = np.empty(2 * nx + 2 * ny)
        _check(self.lib, self.h, self.lib.dense_get_theta(self.h, _dp(buf)))
        return (buf[:nx].reshape(nj, ni + 1), buf[nx:2 * nx].reshape(nj, ni + 1),
                buf[2 * nx:2 * nx + ny].reshape(ni, nj + 1), buf[2 * nx + ny:].reshape(ni, nj + 1))

    def sync(self):
        _check(self.lib, self.h, self.lib.dense_sync(self.h))


def probe_peaks(hbm_bytes: int = 2 << 30) -> dict:
    """Roofline calibration of the current device: FP64 FMA TFLOP/s and the
    HBM streaming read and copy rates in GB/s (buffers of hbm_bytes)."""
    lib = load_library()
    t = C.c_double(0.0)
    r, c = C.c_double(0.0), C.c_double(0.0)
    _check(lib, None, lib.dense_probe_fp64(C.byref(t)))
    _check(lib, None, lib.dense_probe_hbm(hbm_bytes, C.byref(r), C.byref(c)))
    return dict(fp64_tflops=t.value, hbm_read_gbs=r.value, hbm_copy_gbs=c.value, hbm_bytes=hbm_bytes)


def cell_updates(ni: int, nj: int, steps: int) -> int:
    """One cell-update = one interior cell advanced one full RK step."""
    return ni * nj * steps
