"""Pins the oracle's non-periodic z boundaries (bc_sym_z, boundcond.f90:
3842-3900: 's', 'a', 'a2') and constant vertical gravity (gravity_simple,
gravz_profile='const') against what the mathematics fixes.

* An isothermal hydrostatic atmosphere, lnrho linear in z with slope
  gravz/cs^2 and u = 0, is an exact steady state of the discrete equations:
  the sixth-order first derivative is exact on a linear function and 'a2'
  continues a linear function into the ghost zones.  A wrong sign of the
  gravity term, or 'a2' built as 's', breaks it.
* On a grid whose boundary points are z0 and z0 + Lz (dz = Lz/(nz-1)), the
  ghost rules are the even ('s'), odd ('a') and odd-about-the-boundary-value
  ('a2') continuations, so cos(pi m (z-z0)/Lz), sin(...) and c + sin(...)
  have the exact discrete second-derivative symbol at every point,
  boundary points included: a shear flow u_x(z) decays at nu*k2 everywhere.
* 'a' sets the boundary value to zero before the RHS, so the value after a
  step at the boundary is dt*df there, as the reference leaves it."""
import math

import numpy as np
import pytest

import oracle as O
from test_oracle_stencils import k2_sym

N = (8, 6, 16)
LZ = 2.0


def zgrid(nz=N[2], Lz=LZ, z0=-0.7):
    dz = Lz / (nz - 1)
    return dz, z0, (z0 + dz * np.arange(nz))[:, None, None]


def params(bcz, **kw):
    return O.make_params(N, L=(1.0, 1.5, LZ), bcz=bcz, **kw)


@pytest.mark.parametrize("lmag", [False, True])
def test_isothermal_hydrostatic_atmosphere_is_steady(lmag):
    gz, cs0 = -1.3, 0.8
    dz, z0, Z = zgrid()
    bcz = ["s", "s", "a", "a2"] + (["a", "a", "s"] if lmag else [])
    p = params(bcz, gravz=gz, cs0=cs0, lmagnetic=lmag, eta=1e-3)
    f = np.zeros((O.nvar(p),) + (N[2], N[1], N[0]))
    f[3] = 0.4 + gz / cs0 ** 2 * (Z - z0) * np.ones((N[1], N[0]))
    r = O.rhs(f, p)[0]
    assert np.abs(r).max() < 1e-14
    # the same state without gravity accelerates upward (pressure gradient alone)
    r0 = O.rhs(f, params(bcz, cs0=cs0, lmagnetic=lmag, eta=1e-3))[0]
    np.testing.assert_allclose(r0[2], -gz * np.ones_like(r0[2]), rtol=1e-12)
    # 'a2' on lnrho built as 's' breaks it at the boundaries only
    bad = params(bcz[:3] + ["s"] + bcz[4:], gravz=gz, cs0=cs0, lmagnetic=lmag, eta=1e-3)
    rb = O.rhs(f, bad)[0]
    assert np.abs(rb[2, 3:-3]).max() < 1e-14 and np.abs(rb[2, 0]).max() > 0.1


@pytest.mark.parametrize("code,shape", [("s", "cos"), ("a", "sin"), ("a2", "c+sin")])
@pytest.mark.parametrize("m", [1, 3])
def test_shear_wave_symbol_with_mirror_ghosts(code, shape, m):
    nu = 0.05
    dz, z0, Z = zgrid()
    k = math.pi * m / LZ
    base = {"cos": np.cos(k * (Z - z0)), "sin": np.sin(k * (Z - z0)), "c+sin": 0.3 + np.sin(k * (Z - z0))}[shape]
    p = params([code, "s", "a", "a2"], nu=nu)
    f = np.zeros((4, N[2], N[1], N[0]))
    f[0] = base * np.ones((N[1], N[0]))
    if code == "a":
        f[0][0] = 0.0
        f[0][-1] = 0.0
    r = O.rhs(f, p)[0]
    osc = base - (0.3 if shape == "c+sin" else 0.0)
    expect = -nu * k2_sym(k, dz) * osc * np.ones((N[1], N[0]))
    np.testing.assert_allclose(r[0], expect, rtol=0, atol=1e-13)


def test_ghost_rules_and_boundary_zeroing():
    p = params(["s", "a", "a2", "a2:s"])
    nz = N[2]
    rng = np.random.default_rng(3)
    interior = rng.standard_normal((4, nz, N[1], N[0]))
    f = O.with_ghosts(interior)
    O.boundconds(f, p)
    g = 3
    bot, top = g, g + nz - 1
    for i in (1, 2, 3):
        np.testing.assert_array_equal(f[0, bot - i], f[0, bot + i])
        np.testing.assert_array_equal(f[0, top + i], f[0, top - i])
        np.testing.assert_array_equal(f[1, bot - i], -f[1, bot + i])
        np.testing.assert_array_equal(f[1, top + i], -f[1, top - i])
        np.testing.assert_array_equal(f[2, bot - i], 2 * f[2, bot] - f[2, bot + i])
        np.testing.assert_array_equal(f[2, top + i], f[2, top] + (f[2, top] - f[2, top - i]))
        np.testing.assert_array_equal(f[3, bot - i], 2 * f[3, bot] - f[3, bot + i])   # 'a2' below
        np.testing.assert_array_equal(f[3, top + i], f[3, top - i])                    # 's' above
    assert np.all(f[1, bot] == 0.0) and np.all(f[1, top] == 0.0)
    np.testing.assert_array_equal(f[0, bot, g:-g, g:-g], interior[0, 0])


def test_antisymmetric_boundary_value_after_a_step():
    """'a' zeroes the boundary value before the RHS; after one Euler-like
    substep sequence the boundary holds only the RK increments."""
    p = params(["s", "s", "a", "a2"], nu=0.02, dt=1e-3)
    dz, z0, Z = zgrid()
    f = np.zeros((4, N[2], N[1], N[0]))
    f[2] = 0.1 * np.cos(math.pi * (Z - z0) / LZ) * np.ones((N[1], N[0]))   # nonzero at the boundaries
    out = O.advance(f, p, 1)[0]
    # u_z = 0 at the boundaries before every RHS; the increments there are
    # small (the ghosts are odd, so the viscous term vanishes at the boundary)
    assert np.abs(out[2, 0]).max() < 1e-3 * np.abs(f[2, 0]).max()
    assert np.abs(out[2, -1]).max() < 1e-3 * np.abs(f[2, -1]).max()
    assert np.abs(out[2, N[2] // 2 - 3]).max() > 0.5 * np.abs(f[2, N[2] // 2 - 3]).max()
