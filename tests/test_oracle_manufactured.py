"""Pins the oracle's full nonlinear RHS assembly by a manufactured solution.

Smooth periodic fields (u, ln rho, [s], [A]) of low wavenumber are given in
closed form; sympy differentiates them and evaluates the continuous
equations of the path as the physics states them:

  d u/dt     = -(u.grad)u - grad p / rho + (1/rho) div(2 rho nu S) + (1/rho) J x B
  d lnrho/dt = -u.grad lnrho - div u
  d s/dt     = -u.grad s + (1/(rho T)) [div(K grad T) + 2 rho nu S^2 + eta mu0 J^2],
               K = rho cp chi
  d A/dt     = u x B + eta lap A,     B = curl A,  J = curl B / mu0

with p = rho cs^2 / gamma, cs^2 = cs0^2 exp(gamma s/cp + (gamma-1)(lnrho - lnrho0))
(isothermal: p = rho cs0^2).  The viscous and pressure forces are written
here in divergence form, not in the S.grad(lnrho) / (glnrho+glnTT) forms the
reference code uses, so a dropped term, a wrong sign or coefficient, or a
transposed index in the oracle shows up as an O(1) error that does not
converge.  The oracle's 6th-order discrete RHS must approach the continuous
one with error ratio ~2^6 when the grid is doubled.
"""
import math

import numpy as np
import pytest

import oracle as O

sp = pytest.importorskip("sympy")

X, Y, Z = sp.symbols("x y z", real=True)
COORDS = (X, Y, Z)


def _fields(lmag, lent):
    u = [0.30 * sp.sin(Y + 0.3) + 0.10 * sp.cos(2 * Z),
         0.20 * sp.sin(Z - 0.7) + 0.15 * sp.cos(X + 2 * Y),
         0.25 * sp.cos(X + 0.2) - 0.10 * sp.sin(Y - Z)]
    lnrho = 0.20 * sp.sin(X) * sp.cos(Y) + 0.10 * sp.sin(Z + 0.4)
    s = 0.15 * sp.cos(X + Y) + 0.05 * sp.sin(2 * Z - 0.1) if lent else None
    A = None
    if lmag:
        A = [0.20 * sp.sin(Z + 0.5) + 0.05 * sp.cos(X - Y),
             0.15 * sp.cos(X - 0.4) + 0.10 * sp.sin(2 * Z),
             0.10 * sp.sin(Y + 0.9) + 0.12 * sp.cos(X + Z)]
    return u, lnrho, s, A


def grad(f):
    return [sp.diff(f, c) for c in COORDS]


def div(v):
    return sum(sp.diff(v[i], COORDS[i]) for i in range(3))


def curl(v):
    return [sp.diff(v[2], Y) - sp.diff(v[1], Z),
            sp.diff(v[0], Z) - sp.diff(v[2], X),
            sp.diff(v[1], X) - sp.diff(v[0], Y)]


def cross(a, b):
    return [a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]]


def lap(f):
    return sum(sp.diff(f, c, 2) for c in COORDS)


def continuous_rhs(lmag, lent, nu, eta, chi, cs0, gamma, cp, rho0, mu0):
    u, lnrho, s, A = _fields(lmag, lent)
    rho = sp.exp(lnrho)
    divu = div(u)
    S = [[sp.Rational(1, 2) * (sp.diff(u[i], COORDS[j]) + sp.diff(u[j], COORDS[i]))
          - (sp.Rational(1, 3) * divu if i == j else 0) for j in range(3)] for i in range(3)]
    # pressure
    if lent:
        cs2 = cs0 ** 2 * sp.exp(gamma * s / cp + (gamma - 1) * (lnrho - math.log(rho0)))
        p = rho * cs2 / gamma
        TT = cs2 / (cp * (gamma - 1))
    else:
        p = rho * cs0 ** 2
        TT = None
    rhs_u = []
    for i in range(3):
        adv = -sum(u[j] * sp.diff(u[i], COORDS[j]) for j in range(3))
        fp = -sp.diff(p, COORDS[i]) / rho
        fv = sum(sp.diff(2 * rho * nu * S[i][j], COORDS[j]) for j in range(3)) / rho
        rhs_u.append(adv + fp + fv)
    if lmag:
        B = curl(A)
        J = [c / mu0 for c in curl(B)]
        jxb = cross(J, B)
        for i in range(3):
            rhs_u[i] = rhs_u[i] + jxb[i] / rho
        uxb = cross(u, B)
        rhs_A = [uxb[i] + eta * lap(A[i]) for i in range(3)]
    rhs_lnrho = -sum(u[j] * sp.diff(lnrho, COORDS[j]) for j in range(3)) - divu
    out = rhs_u + [rhs_lnrho]
    if lent:
        S2 = sum(S[i][j] ** 2 for i in range(3) for j in range(3))
        K = rho * cp * chi
        heat = div([K * g for g in grad(TT)]) + 2 * rho * nu * S2
        if lmag:
            heat = heat + eta * mu0 * sum(J[i] ** 2 for i in range(3))
        rhs_s = -sum(u[j] * sp.diff(s, COORDS[j]) for j in range(3)) + heat / (rho * TT)
        out.append(rhs_s)
    if lmag:
        out += rhs_A
    fields = [u[0], u[1], u[2], lnrho] + ([s] if lent else []) + (A if lmag else [])
    return fields, out


def sample(exprs, n):
    h = 2 * math.pi / n
    x = h * np.arange(n)
    Zg, Yg, Xg = np.meshgrid(x, x, x, indexing="ij")
    vals = []
    for e in exprs:
        fn = sp.lambdify((X, Y, Z), e, "numpy")
        v = np.asarray(fn(Xg, Yg, Zg), dtype=np.float64)
        vals.append(np.broadcast_to(v, Xg.shape).copy())
    return np.array(vals)


@pytest.mark.parametrize("lmag,lent", [(False, False), (True, False), (True, True)])
def test_rhs_converges_to_continuous_equations(lmag, lent):
    par = dict(nu=0.03, eta=0.02 if lmag else 0.0, chi=0.04 if lent else 0.0, cs0=1.0,
               gamma=5.0 / 3.0 if lent else 1.0, cp=1.0, rho0=1.0, mu0=1.0)
    fields, rhs = continuous_rhs(lmag, lent, **par)
    errs = []
    for n in (16, 32):
        f = sample(fields, n)
        exact = sample(rhs, n)
        p = O.make_params(n, lmagnetic=lmag, lentropy=lent, **par)
        got, _, _ = O.rhs(f, p)
        assert got.shape == exact.shape
        errs.append(np.array([np.abs(got[v] - exact[v]).max() / np.abs(exact[v]).max()
                              for v in range(len(fields))]))
    e16, e32 = errs
    # 6th order: halving h divides the error by ~64 for every variable
    assert np.all(e32 < 1e-4), e32
    assert np.all(e16 / e32 > 40), e16 / e32
