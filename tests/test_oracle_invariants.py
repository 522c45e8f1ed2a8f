"""Closed-form checks that pin the oracle's building blocks independently of
the reference's fixtures: properties a dropped term, a wrong sign or a
transposed index would break."""
import numpy as np
import pytest

import pyoracle as po


def _arr(x):
    return np.ascontiguousarray(np.asarray(x, dtype=np.float64))


def _call(fn, lf):
    a = _arr(lf)
    return fn(po._dptr(a))


@pytest.fixture(scope="module")
def lib():
    return po.load()


@pytest.mark.parametrize("a,b", [(1.0, 0.5), (-3.0, 2.0), (10.0, -1.25), (0.0, 1e-3)])
def test_linear_data_all_schemes_reduce_to_upwind5(lib, a, b):
    """On linear data the three smoothness indicators are equal, so WENO-JS and
    WENO-Z weights equal the linear weights (0.1, 0.6, 0.3) and both reproduce
    the upwind5 value, which is the exact interface value a + b/2."""
    lf = [a + b * m for m in (-2, -1, 0, 1, 2)]
    exact = a + 0.5 * b
    for fn in (lib.ora_weno5z_stencil, lib.ora_weno5js_stencil, lib.ora_upwind5_stencil):
        assert _call(fn, lf) == pytest.approx(exact, rel=1e-14, abs=1e-14)
    w = np.zeros(4)
    lib.ora_weno5z_omega_theta(po._dptr(_arr(lf)), po._dptr(w))
    assert w[1:] == pytest.approx([0.1, 0.6, 0.3], rel=1e-14)
    assert w[0] == 1.0  # AdaWENO indicator theta = 1/(1+(sum alpha - 1)^2) = 1


def test_upwind5_exact_on_quartics(lib):
    """The upwind5 interface value from cell averages is exact for polynomials
    of degree <= 4 (5th-order reconstruction)."""
    rng = np.random.default_rng(0)
    for _ in range(20):
        c = rng.normal(size=5)
        P = np.polynomial.Polynomial(c)
        Q = P.integ()
        avg = [Q(m + 0.5) - Q(m - 0.5) for m in (-2, -1, 0, 1, 2)]
        assert _call(lib.ora_upwind5_stencil, avg) == pytest.approx(P(0.5), rel=1e-12, abs=1e-12)


def test_weno_nonoscillatory_at_a_jump(lib):
    """Upwind of a step the non-linear weights put (almost) all weight on the
    smooth candidate stencil (f-2, f-1, f0), while the linear upwind5 value is
    (27 - 3)/60 = 0.4."""
    lf = [0.0, 0.0, 0.0, 1.0, 1.0]
    for fn in (lib.ora_weno5z_stencil, lib.ora_weno5js_stencil):
        h = _call(fn, lf)
        assert 0.0 <= h < 1e-6
    assert _call(lib.ora_upwind5_stencil, lf) == 0.4


def _euler_jacobian_x(u, v, H, gamma):
    g1 = gamma - 1.0
    q2 = 0.5 * (u * u + v * v)
    return np.array([[0.0, 1.0, 0.0, 0.0],
                     [g1 * q2 - u * u, (3.0 - gamma) * u, -g1 * v, g1],
                     [-u * v, v, u, 0.0],
                     [u * (g1 * q2 - H), H - g1 * u * u, -g1 * u * v, gamma * u]])


@pytest.mark.parametrize("seed", range(5))
def test_eigenvectors_invert_and_diagonalise(lib, seed):
    """L R = I and R diag(u-c, u, u, u+c) L = dF/dU for the x Jacobian; the y
    pair diagonalises dG/dU (the x Jacobian with u<->v, momentum components swapped)."""
    rng = np.random.default_rng(seed)
    gamma = 1.4
    u, v = rng.normal(size=2)
    c = 1.0 + rng.random()
    H = c * c / (gamma - 1.0) + 0.5 * (u * u + v * v)
    l = np.zeros(16)
    r = np.zeros(16)
    lib.ora_lr_vec_xi(u, v, H, c, gamma, po._dptr(l), po._dptr(r))
    lm = l.reshape(4, 4, order="F")  # l(comp, char)
    rm = r.reshape(4, 4, order="F")  # r(char, comp)
    Lrows = lm.T  # row k = left eigenvector of char k
    Rcols = rm.T  # column k = right eigenvector of char k
    assert np.allclose(Lrows @ Rcols, np.eye(4), atol=1e-12)
    A = _euler_jacobian_x(u, v, H, gamma)
    assert np.allclose(Rcols @ np.diag([u - c, u, u, u + c]) @ Lrows, A, atol=1e-11)
    lib.ora_lr_vec_eta(u, v, H, c, gamma, po._dptr(l), po._dptr(r))
    Lrows = l.reshape(4, 4, order="F").T
    Rcols = r.reshape(4, 4, order="F").T
    P = np.eye(4)[[0, 2, 1, 3]]  # swap momentum components
    B = P @ _euler_jacobian_x(v, u, H, gamma) @ P
    assert np.allclose(Lrows @ Rcols, np.eye(4), atol=1e-12)
    assert np.allclose(Rcols @ np.diag([v - c, v, v, v + c]) @ Lrows, B, atol=1e-11)


def _periodic_params(n, splitting, recon, cfl=0.4):
    p = po.problem_params(1, n, n, splitting=splitting, recon=recon, cfl=cfl)
    for s in range(4):
        p.bc[s] = po.BC_PERIODIC
    return p


def _uniform(n, gamma=1.4, rho=1.3, u=0.7, v=-0.4, p=2.0):
    c = np.zeros((4, n + 8, n + 8))
    c[0] = rho
    c[1] = rho * u
    c[2] = rho * v
    c[3] = p / (gamma - 1.0) + 0.5 * rho * (u * u + v * v)
    return c


@pytest.mark.parametrize("sp,rc", [(101, 255), (102, 253), (103, 251), (104, 255), (105, 255)])
def test_free_stream_preserved(sp, rc):
    """A uniform state is a steady solution: F+ + F- = F and every stencil
    reproduces constants, so the state stays uniform."""
    n = 16
    o = po.Oracle(_periodic_params(n, sp, rc))
    c0 = _uniform(n)
    o.set_cons(c0)
    o.step(3)
    got = o.interior()
    ref = c0[:, 4:-4, 4:-4]
    assert np.max(np.abs(got - ref) / np.abs(ref).max(axis=(1, 2), keepdims=True)) < 1e-13


@pytest.mark.parametrize("sp,rc", [(101, 253), (103, 255), (104, 255), (105, 255)])
def test_conservation_periodic(sp, rc):
    """With periodic boundaries the face fluxes telescope: total mass,
    momentum and energy are conserved to round-off."""
    n = 32
    o = po.Oracle(_periodic_params(n, sp, rc))
    src = po.Oracle(po.problem_params(1, n, n))
    src.init_state(1)  # smooth vortex (reference's IVC), made periodic below
    c = src.cons()
    o.set_cons(c)
    before = o.interior().sum(axis=(1, 2))
    o.step(5)
    after = o.interior().sum(axis=(1, 2))
    scale = np.abs(o.interior()).sum(axis=(1, 2))
    assert np.all(np.abs(after - before) <= 1e-12 * scale + 1e-14)


def _euler_flux(rho, u, v, p, gamma, nrm):
    E = p / (gamma - 1.0) + 0.5 * rho * (u * u + v * v)
    un = u if nrm == 2 else v
    f = np.array([rho * un, rho * un * u, rho * un * v, un * (E + p)])
    f[nrm - 1] += p
    return f, np.array([rho, rho * u, rho * v, E])


def _split(lib, kind, rho, u, v, p, gamma, nrm, a=None):
    c = np.sqrt(gamma * p / rho)
    fp, fn = np.zeros(4), np.zeros(4)
    dp = lambda x: x.ctypes.data_as(po.C.POINTER(po.C.c_double))  # noqa: E731
    if kind == "lf":
        _, U = _euler_flux(rho, u, v, p, gamma, nrm)
        un = u if nrm == 2 else v
        lib.ora_lf_split(un, p, abs(un) + c if a is None else a, dp(U), nrm, dp(fp), dp(fn))
    else:
        lib.ora_sw_split(gamma, rho, u, v, c, nrm, dp(fp), dp(fn))
    return fp, fn


@pytest.mark.parametrize("seed", range(4))
def test_flux_splits_add_up_to_the_euler_flux(lib, seed):
    """Lax-Friedrichs (flux_splitting_LLF.f90:37-47, GLF the same with a global
    speed) and Steger-Warming (flux_splitting_SW.f90:36-161): F+ + F- = F(U)
    in both directions; for LF also F+ - F- = alpha U, and Steger-Warming
    sends a supersonic flow almost entirely to F+ (lambda- = (lambda - sqrt(
    lambda^2 + eps^2))/2 is O(eps^2/lambda) with eps = 0.125)."""
    rng = np.random.default_rng(seed)
    gamma = 1.4
    for _ in range(20):
        rho, p = rng.uniform(0.2, 5.0, 2)
        u, v = rng.uniform(-3.0, 3.0, 2)
        for nrm in (2, 3):
            F, U = _euler_flux(rho, u, v, p, gamma, nrm)
            for kind in ("lf", "sw"):
                fp, fn = _split(lib, kind, rho, u, v, p, gamma, nrm)
                assert np.allclose(fp + fn, F, rtol=1e-13, atol=1e-13 * np.abs(F).max()), (kind, nrm)
            a = rng.uniform(0.0, 10.0)
            fp, fn = _split(lib, "lf", rho, u, v, p, gamma, nrm, a=a)
            assert np.allclose(fp - fn, a * U, rtol=1e-13, atol=1e-13 * np.abs(a * U).max())
    rho, p = 1.0, 1.0
    c = np.sqrt(gamma)
    F, _ = _euler_flux(rho, 8.0 * c, 0.3, p, gamma, 2)
    fp, fn = _split(lib, "sw", rho, 8.0 * c, 0.3, p, gamma, 2)
    assert np.abs(fn).max() < 1e-3 * np.abs(F).max()


def test_adaweno_threshold_one_is_the_characteristic_scheme():
    """theta = 1/(1 + (sum alpha - 1)^2) <= 1 (weno_mod.f90:94), so with
    delta >= 1 no face takes the component-wise branch (theta > delta,
    scheme_weno5th_SC.f90:110) and switch 105 is switch 104 bit for bit; with
    the default delta = 0.5 it is not."""
    outs = {}
    for sp, delta in ((104, 0.5), (105, 1.0), (105, 0.5)):
        p = po.problem_params(5, 64, 16, splitting=sp, recon=255, cfl=0.6, delta=delta)
        o = po.Oracle(p)
        o.init_state(5)
        o.step(8)
        outs[(sp, delta)] = o.interior()
        o.close()
    assert np.array_equal(outs[(105, 1.0)], outs[(104, 0.5)])
    assert not np.array_equal(outs[(105, 0.5)], outs[(104, 0.5)])
