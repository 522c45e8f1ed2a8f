"""dense_run_noalloc: the allocation-free variant of dense_run.

Its scratch memory (device clock, launch counters, two stage buffers) lives in
a file-scope device array of the library, so the entry point calls no
allocator.  Checked through the C ABI against dense_run on the same inputs,
bit for bit, on two calls in a row: a double Mach 960x240 block (characteristic
WENO-Z, its time-dependent top boundary) and then a smaller Rayleigh-Taylor
256x128 block with gravity and AdaWENO, whose stage buffers reuse the start
of the array the first call filled.  Also: the size limit is enforced with an
error, and a state with a row pitch above ni is stepped in place.
"""
import ctypes as C
import importlib

import numpy as np
import pytest

from conftest import dense

pytestmark = pytest.mark.gpu


def _lib():
    D = dense()
    lib = D.load_library()
    lib.dense_run_noalloc.argtypes = [C.POINTER(D.Cfg), C.c_void_p, C.c_int, C.c_int, C.c_double, C.c_void_p,
                                      C.POINTER(C.c_double), C.POINTER(C.c_double)]
    lib.dense_run_noalloc.restype = C.c_int
    return lib


def _noalloc(cfg, cons, steps, ld=None, t0=0.0):
    """Advance the interior state cons (4, nj, ni) in device memory held by
    torch; returns (state, t, dt of the last step)."""
    import torch
    lib = _lib()
    ni, nj = cfg.ni, cfg.nj
    ld = ld or ni
    u = torch.zeros((4, nj, ld), dtype=torch.float64, device="cuda")
    u[:, :, :ni] = torch.from_numpy(np.ascontiguousarray(cons)).to("cuda")
    torch.cuda.synchronize()
    t, dt = C.c_double(0.0), C.c_double(0.0)
    stream = torch.cuda.current_stream().cuda_stream
    rc = lib.dense_run_noalloc(C.byref(cfg), C.c_void_p(u.data_ptr()), ld, steps, t0, C.c_void_p(stream),
                               C.byref(t), C.byref(dt))
    if rc != 0:
        raise dense().DenseError(rc, lib.dense_last_error(None).decode())
    torch.cuda.synchronize()
    return u[:, :, :ni].cpu().numpy(), t.value, dt.value


def _with_run(cfg, cons, steps):
    D = dense()
    with D.Solver(cfg) as s:
        s.set_state(cons)
        taken, dts = s.run(steps, want_dts=True)
        assert taken == steps
        return s.interior(), s.current_time, float(dts[-1])


def test_noalloc_equals_dense_run_on_two_calls_the_second_smaller():
    D = dense()
    bench = importlib.import_module("bench")
    cases = [
        ("double Mach 960x240, 104", D.make_cfg(960, 240, benchmark=5, splitting=104, recon=255, cfl=0.6),
         bench.dmr_initial_state(960, 240), 12),
        ("Rayleigh-Taylor 256x128, 105 + gravity", D.make_cfg(256, 128, benchmark=3, splitting=105, recon=255,
                                                             cfl=0.5), bench.rt_initial_state(256, 128), 9),
    ]
    for name, cfg, cons, steps in cases:
        print(f"no-alloc vs dense_run: {name}, {steps} steps", flush=True)
        got, t, dt = _noalloc(cfg, cons, steps)
        ref, t_ref, dt_ref = _with_run(cfg, cons, steps)
        assert np.array_equal(got, ref), f"{name}: max |diff| {np.max(np.abs(got - ref))}"
        assert t == t_ref and dt == dt_ref


def test_noalloc_pitched_state_and_limits():
    D = dense()
    B = importlib.import_module("2dense_amd.batch")
    cons, h = B.vortex_state(100, 3.0, 4.0, 6.0)
    cfg = D.make_cfg(100, 100, benchmark=0, splitting=101, recon=253, cfl=0.5, gamma=1.4, dx=h, dy=h,
                     bc=[D.BC_PERIODIC] * 4)
    print("no-alloc: 100x100 vortex, row pitch 128", flush=True)
    got, t, _ = _noalloc(cfg, cons, 7, ld=128)
    ref, t_ref, _ = _with_run(cfg, cons, 7)
    assert np.array_equal(got, ref) and t == t_ref
    lib = _lib()
    big = D.make_cfg(2048, 1024, benchmark=5, splitting=104, recon=255, cfl=0.6)
    rc = lib.dense_run_noalloc(C.byref(big), C.c_void_p(16), 2048, 1, 0.0, None, None, None)
    assert rc == -1 and b"allocation-free limit" in lib.dense_last_error(None)
    rc = lib.dense_run_noalloc(C.byref(cfg), C.c_void_p(16), 99, 1, 0.0, None, None, None)
    assert rc == -1
