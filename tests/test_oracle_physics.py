"""Pins the oracle's RHS assembly, time step and diagnostics against what the
physics fixes: linear theory with the exact semi-discrete symbols of the
stencils, force balances whose sign is physical, closed-form Courant and
rms values, invariances, and the reference's own printed known answers
(samples/gputest/cpu-test.out)."""
import math

import numpy as np
import pytest

import oracle as O
from test_oracle_stencils import grid, k1_sym, k2_sym

TWO_PI = 2 * math.pi


def zeros(p):
    return np.zeros((O.nvar(p), p.nz, p.ny, p.nx))


def rk3_factor(z):
    return 1 + z + z * z / 2 + z ** 3 / 6


# ---------------------------------------------------------------- RK3 + viscosity
@pytest.mark.parametrize("k", [1, 3])
def test_rk3_amplification_viscous_shear_wave(k):
    n = 16
    h, X, Y, Z = grid(n)
    nu, dt = 0.05, 0.01
    p = O.make_params(n, nu=nu, dt=dt)
    f = zeros(p)
    f[0] = 0.3 * np.sin(k * Y) * np.ones_like(X)   # u_x(y): divergence free, u.grad u = 0
    out, t, _, _ = O.advance(f, p, 1)
    z = -nu * k2_sym(k, h) * dt
    np.testing.assert_allclose(out[0], f[0] * rk3_factor(z), rtol=0, atol=1e-15)
    assert np.abs(out[1:]).max() < 1e-17
    assert abs(t - dt) < 1e-17
    # two steps: factor squared (the 2N scheme's alpha/beta reproduce R(z) exactly)
    out2, t2, _, _ = O.advance(f, p, 2)
    np.testing.assert_allclose(out2[0], f[0] * rk3_factor(z) ** 2, rtol=0, atol=1e-15)
    assert abs(t2 - 2 * dt) < 1e-16


# ---------------------------------------------------------------- sound waves
def _linear_sound_step(a0, b0, k1, k2, cs2, nu, dt):
    # d/dt (a, b) = M (a, b): lnrho = a cos kx, u_x = b sin kx
    M = np.array([[0.0, -k1], [cs2 * k1, -(4.0 / 3.0) * nu * k2]])
    A = dt * M
    R = np.eye(2) + A + A @ A / 2 + A @ A @ A / 6
    return R @ np.array([a0, b0])


@pytest.mark.parametrize("lentropy", [False, True])
def test_linear_sound_wave(lentropy):
    # isothermal: cs2 = cs20; with entropy (gamma=5/3, ss=0): the adiabatic
    # response gives the same linear system with cs2 = cs20 (entropy.f90 fpres)
    n, k = 16, 2
    h, X, Y, Z = grid(n)
    gamma = 5.0 / 3.0 if lentropy else 1.0
    p = O.make_params(n, lentropy=lentropy, gamma=gamma, nu=0.02, dt=2e-3)
    eps = 1e-7
    res = []
    for s in (+1, -1):
        f = zeros(p)
        f[3] = s * eps * np.cos(k * X) * np.ones_like(Y * Z)
        f[0] = s * 0.5 * eps * np.sin(k * X) * np.ones_like(Y * Z)
        out, _, _, _ = O.advance(f, p, 1)
        res.append(out)
    odd = 0.5 * (res[0] - res[1])          # removes the O(eps^2) nonlinear terms
    a, b = _linear_sound_step(eps, 0.5 * eps, k1_sym(k, h), k2_sym(k, h), 1.0, p.nu, p.dt)
    np.testing.assert_allclose(odd[3], a * np.cos(k * X) * np.ones_like(Y * Z), rtol=0, atol=1e-18)
    np.testing.assert_allclose(odd[0], b * np.sin(k * X) * np.ones_like(Y * Z), rtol=0, atol=1e-18)


def test_advection_term_sign():
    # u_x = U sin x, everything else zero: du_x/dt = -u_x D1x u_x + nu*(4/3) D2x u_x
    n = 16
    h, X, Y, Z = grid(n)
    p = O.make_params(n, nu=0.01)
    f = zeros(p)
    U = 0.2
    f[0] = U * np.sin(X) * np.ones_like(Y * Z)
    r, _, _ = O.rhs(f, p)
    k1, k2 = k1_sym(1, h), k2_sym(1, h)
    expect = -U * np.sin(X) * U * k1 * np.cos(X) - p.nu * (4.0 / 3.0) * k2 * U * np.sin(X)
    np.testing.assert_allclose(r[0], expect * np.ones_like(Y * Z), rtol=0, atol=1e-15)
    # continuity: dlnrho/dt = -div u
    np.testing.assert_allclose(r[3], -U * k1 * np.cos(X) * np.ones_like(Y * Z), rtol=0, atol=1e-15)


# ---------------------------------------------------------------- magnetic
def _helical(p, a, k, X, Y, Z):
    f = zeros(p)
    ia = O.iax(p)
    f[ia] = a * np.sin(k * Z) * np.ones_like(X * Y)
    f[ia + 1] = a * np.cos(k * Z) * np.ones_like(X * Y)
    return f


def test_force_free_field_and_resistive_decay():
    # A = a (sin kz, cos kz, 0): B = k1 A, J = k2 A, so J x B = 0 and
    # dA/dt = eta del2 A = -eta k2 A (u = 0): each step multiplies A by R(-eta k2 dt)
    n, k = 16, 2
    h, X, Y, Z = grid(n)
    eta, dt = 0.03, 5e-3
    p = O.make_params(n, lmagnetic=True, nu=0.01, eta=eta, dt=dt)
    f = _helical(p, 0.1, k, X, Y, Z)
    r, _, diag = O.rhs(f, p)
    assert np.abs(r[:4]).max() < 1e-14 * (k * k * 0.1) ** 2     # no Lorentz force, no flow
    np.testing.assert_allclose(r[4:6], -eta * k2_sym(k, h) * f[4:6], rtol=0, atol=1e-16)
    # brms = |B| = k1 a (uniform magnitude)
    assert abs(diag[2] - k1_sym(k, h) * 0.1) < 1e-15
    out, _, _, _ = O.advance(f, p, 3)
    np.testing.assert_allclose(out[4:6], f[4:6] * rk3_factor(-eta * k2_sym(k, h) * dt) ** 3, rtol=0, atol=1e-16)


def test_magnetic_pressure_force():
    # A_z = a cos kx: B_y = a k1 sin kx, J_z = a k2 cos kx,
    # (J x B)_x = -J_z B_y = -a^2 k1 k2 sin cos  (pushes from high to low B^2)
    n, k = 16, 1
    h, X, Y, Z = grid(n)
    p = O.make_params(n, lmagnetic=True, nu=0.0, eta=0.0)
    f = zeros(p)
    a = 0.05
    f[6] = a * np.cos(k * X) * np.ones_like(Y * Z)
    f[3] = 0.2                                                 # lnrho: rho1 = exp(-0.2)
    r, _, _ = O.rhs(f, p)
    expect = -a * a * k1_sym(k, h) * k2_sym(k, h) * np.sin(k * X) * np.cos(k * X) * math.exp(-0.2)
    np.testing.assert_allclose(r[0], expect * np.ones_like(Y * Z), rtol=0, atol=1e-13 * np.abs(expect).max())
    assert np.abs(r[1:3]).max() < 1e-15 * np.abs(expect).max()


def test_induction_advects_potential_downstream():
    # uniform u = U x, A_z = a sin kx:  dA_z/dt = (u x B)_z = -U D1x A_z
    n, k = 16, 2
    h, X, Y, Z = grid(n)
    p = O.make_params(n, lmagnetic=True, nu=0.0, eta=0.0)
    f = zeros(p)
    U, a = 0.7, 1e-3
    f[0] = U
    f[6] = a * np.sin(k * X) * np.ones_like(Y * Z)
    r, _, _ = O.rhs(f, p)
    np.testing.assert_allclose(r[6], -U * a * k1_sym(k, h) * np.cos(k * X) * np.ones_like(Y * Z),
                               rtol=0, atol=1e-17)


# ---------------------------------------------------------------- Courant time step
def test_courant_dt1_uniform_flow_isothermal():
    n = 16
    p = O.make_params(n, nu=0.02, dt=0.0)
    f = zeros(p)
    U = (0.3, -0.2, 0.1)
    for i in range(3):
        f[i] = U[i]
    _, dt1, _ = O.rhs(f, p, want_dt=True)
    dx_1 = n / TWO_PI
    dxyz_2 = 3 * dx_1 ** 2
    adv = sum(abs(u) for u in U) * dx_1 + math.sqrt(1.0 * dxyz_2)
    expect = math.sqrt((adv / 0.9) ** 2 + (0.02 * dxyz_2 / 0.25) ** 2)
    assert abs(dt1 - expect) < 1e-13 * expect
    # the step the reference takes is 1/dt1 (sub.f90 set_dt)
    _, _, dt, _ = O.advance(f, p, 1)
    assert abs(dt - 1 / expect) < 1e-13 / expect


def test_courant_dt1_alfven_and_resistive():
    n, k, a = 16, 2, 0.3
    h, X, Y, Z = grid(n)
    p = O.make_params(n, lmagnetic=True, nu=0.01, eta=0.04, dt=0.0, cdt=0.4)
    f = _helical(p, a, k, X, Y, Z)
    _, dt1, _ = O.rhs(f, p, want_dt=True)
    dx_1 = n / TWO_PI
    dxyz_2 = 3 * dx_1 ** 2
    B = k1_sym(k, h) * a
    va2 = (B * dx_1) ** 2                          # |B|^2 sum over the two in-plane components
    adv = math.sqrt(va2 + dxyz_2)
    expect = math.sqrt((adv / 0.4) ** 2 + (0.04 * dxyz_2 / 0.25) ** 2)
    assert abs(dt1 - expect) < 1e-12 * expect


# ---------------------------------------------------------------- diagnostics
def test_urms_umax_closed_form():
    n = 12
    h, X, Y, Z = grid(n)
    p = O.make_params(n)
    f = zeros(p)
    f[0] = np.sin(Y) * np.ones_like(X * Z)
    f[1] = np.sin(Z) * np.ones_like(X * Y)
    f[2] = np.sin(X) * np.ones_like(Y * Z)
    _, _, (urms, umax, brms) = O.rhs(f, p)
    assert abs(urms - math.sqrt(1.5)) < 1e-14
    u2 = f[0] ** 2 + f[1] ** 2 + f[2] ** 2
    assert abs(umax - math.sqrt(u2.max())) < 1e-15
    assert brms == 0.0


# ---------------------------------------------------------------- invariances
def _random_state(p, seed, amp=1e-2):
    rng = np.random.default_rng(seed)
    f = amp * rng.standard_normal((O.nvar(p), p.nz, p.ny, p.nx))
    return f


@pytest.mark.parametrize("lmag,lent", [(False, False), (True, False), (True, True)])
def test_translation_invariance_is_bitwise(lmag, lent):
    n = 8
    p = O.make_params(n, lmagnetic=lmag, lentropy=lent, gamma=5 / 3 if lent else 1.0,
                      nu=0.01, eta=0.01, chi=0.01)
    f = _random_state(p, 3)
    out, _, _, _ = O.advance(f, p, 2)
    sh, _, _, _ = O.advance(np.roll(f, (1, 2, 3), axis=(1, 2, 3)), p, 2)
    np.testing.assert_array_equal(np.roll(out, (1, 2, 3), axis=(1, 2, 3)), sh)


def _rotate(f, p):
    """Cyclic relabelling of axes x->y->z->x applied to fields and components."""
    g = np.transpose(f, (0, 2, 3, 1)).copy()      # new (z,y,x) index = old (y,x,z)? see below
    return g


@pytest.mark.parametrize("lmag,lent", [(False, False), (True, False), (True, True)])
def test_cyclic_axis_permutation_symmetry(lmag, lent):
    # relabel (x,y,z) -> (y,z,x): the new field at (x',y',z') = old at
    # (x=z', y=x', z=y'), vector components permute the same way.  The
    # equations are invariant; the stencils are applied in other directions,
    # so results agree to round-off.
    n = 8
    p = O.make_params(n, lmagnetic=lmag, lentropy=lent, gamma=5 / 3 if lent else 1.0,
                      nu=0.01, eta=0.02, chi=0.015)
    f = _random_state(p, 5, amp=0.05)

    def perm(a):
        # array axes (z, y, x); new[z', y', x'] = old[z=y', y=x', x=z']
        b = np.transpose(a, (0, 2, 3, 1))
        out = b.copy()
        for base in [0] + ([O.iax(p)] if lmag else []):
            # new component x' = old z, y' = old x, z' = old y
            out[base + 0], out[base + 1], out[base + 2] = b[base + 2], b[base + 0], b[base + 1]
        return out

    r1, _, d1 = O.rhs(perm(f), p)
    r0, _, d0 = O.rhs(f, p)
    np.testing.assert_allclose(r1, perm(r0), rtol=0, atol=1e-13 * np.abs(r0).max())
    np.testing.assert_allclose(d1, d0, rtol=1e-13)


# ---------------------------------------------------------------- entropy / EOS
GAMMA = 5.0 / 3.0


def test_eos_constants_match_reference_log():
    # samples/gputest/cpu-test.out:243: "cp, lnTT0, cs0, pp0, Rgas= 1., 0.40546510810816422, 1., ..."
    p = O.make_params(8, lentropy=True, gamma=GAMMA)
    c = O.eos_constants(p)
    assert c["lnTT0"] == 0.40546510810816422
    assert abs(c["cv1"] - GAMMA) < 3e-16 and abs(c["gamma_m1"] - 2 / 3) < 3e-16


def test_cs2_and_TT1_match_reference_log_point():
    # cpu-test.out:270: "denergy_dt: lnTT,cs2,cp1= 1.2254651081081642, 2.2704998375324057"
    # i.e. cv1*ss + gamma_m1*(lnrho-lnrho0) = 0.82.  Take lnrho=0, ss=0.82/cv1.
    n = 8
    h, X, Y, Z = grid(n)
    nu = 0.01
    p = O.make_params(n, lentropy=True, gamma=GAMMA, nu=nu, chi=0.0, dt=0.0)
    f = zeros(p)
    f[4] = 0.82 / GAMMA
    # cs2 enters dt1 through advec_cs2 = cs2*dxyz_2 (u = 0)
    _, dt1, _ = O.rhs(f, p, want_dt=True)
    dxyz_2 = 3 * (n / TWO_PI) ** 2
    cs2 = ((dt1 ** 2 - (nu * dxyz_2 / 0.25) ** 2) ** 0.5 * 0.9) ** 2 / dxyz_2
    assert abs(cs2 - 2.2704998375324057) < 2e-13
    # TT1 = exp(-lnTT) enters ds/dt through viscous heating of a shear flow:
    # ds/dt = TT1 * 2 nu S^2 = TT1 * nu (U k1 cos ky)^2
    U = 0.1
    f[0] = U * np.sin(Y) * np.ones_like(X * Z)
    r, _, _ = O.rhs(f, p)
    k1 = k1_sym(1, h)
    heat = nu * (U * k1 * np.cos(Y)) ** 2 * np.ones_like(X * Z)
    TT1 = r[4][heat > 1e-6 * heat.max()] / heat[heat > 1e-6 * heat.max()]
    np.testing.assert_allclose(TT1, math.exp(-1.2254651081081642), rtol=1e-12)


def test_entropy_diffusion_rate_is_gamma_chi():
    # linear part of thdiff: chi*cp*cv1*del2 ss = gamma*chi*D2 ss
    n, k = 16, 2
    h, X, Y, Z = grid(n)
    chi = 0.02
    p = O.make_params(n, lentropy=True, gamma=GAMMA, nu=0.0, chi=chi)
    eps = 1e-6
    rs = []
    for s in (1, -1):
        f = zeros(p)
        f[4] = s * eps * np.sin(k * Z) * np.ones_like(X * Y)
        rs.append(O.rhs(f, p)[0])
    odd = 0.5 * (rs[0] - rs[1])
    expect = -GAMMA * chi * k2_sym(k, h) * eps * np.sin(k * Z) * np.ones_like(X * Y)
    np.testing.assert_allclose(odd[4], expect, rtol=0, atol=1e-19)
    # pressure response: du_z/dt = -cs2*cv1*gss/gamma = -cs20 * D1z ss (linear)
    # (the odd part keeps O(eps^3) terms of exp(cv1*ss): relative ~gamma^2 eps^2)
    np.testing.assert_allclose(odd[2], -k1_sym(k, h) * eps * np.cos(k * Z) * np.ones_like(X * Y),
                               rtol=0, atol=1e-11 * k1_sym(k, h) * eps)


def test_ohmic_heating():
    # force-free A: rho T ds/dt = eta mu0 J^2 with J = k2 A
    n, k, a = 16, 2, 0.2
    h, X, Y, Z = grid(n)
    eta = 0.03
    p = O.make_params(n, lmagnetic=True, lentropy=True, gamma=GAMMA, nu=0.0, eta=eta, chi=0.0)
    f = _helical(p, a, k, X, Y, Z)
    r, _, _ = O.rhs(f, p)
    TT1 = math.exp(-O.eos_constants(p)["lnTT0"])
    expect = eta * (k2_sym(k, h) * a) ** 2 * TT1
    np.testing.assert_allclose(r[4], expect * np.ones_like(r[4]), rtol=1e-12)
