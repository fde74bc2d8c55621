"""Pins the oracle's samples/conv-slab physics against what the mathematics
and the physics fix:

* Coriolis force (coriolis_cartesian, hydro.f90:6339-6413) equals
  -2 Omega x u with Omega = Omega (-sin theta, 0, cos theta) (a cross product
  done by numpy), and does no work.
* K-profile heat conduction (calc_heatcond, entropy.f90:5913-6151) with a
  uniform K = hcond0 at uniform density equals chi-const conduction with
  chi = K/(rho cp) (pinned elsewhere), Courant time step included; with the
  tanh profile of get_gravz_heatcond on a state with ln T linear in z it
  gives rho1 K(z) g (g + dlnK/dz), K(z) and its log-derivative taken from a
  numerically differentiated numpy profile.
* Cooling towards cs2cool and bottom heating (get_heat_cool_gravz): the
  entropy source is (1/(rho T)) (heat - cool*prof*(cs2 - cs2cool)/cs2cool)
  with T = cs2/((gamma-1) cp), negative where cs2 > cs2cool; the bottom
  heating profile integrates to the luminosity over the box.
* 'c1' (bc_ss_flux): for a state whose conductive flux -K dT/dz equals Fbot
  at the boundary and whose lnrho and ss are linear in z, the ghost zones
  are the linear continuation.
* 'cT' (bc_ss_temp_z): cs2 at the boundary equals cs2top and ln cs2 is odd
  about ln cs2top across the boundary."""
import math

import numpy as np
import pytest

import oracle as O

GAMMA = 1.6666666
N = (8, 6, 24)
L = (1.0, 1.0, 2.0)
Z0 = -0.68


def conv_layers(**kw):
    base = dict(hcond0=8e-3, mpoly0=1.0, mpoly1=3.0, mpoly2=0.0, widthss=0.05, z1=0.0, z2=1.0,
                wheat=0.1, cool=15.0, wcool=0.2, cs2cool=0.333333)
    base.update(kw)
    return O.layer_defaults(GAMMA, 1.0, -1.0, "c1:cT", **base)


def zgrid(nz=N[2], Lz=L[2], z0=Z0):
    dz = Lz / (nz - 1)
    return dz, z0 + dz * np.arange(nz)


def ent_state(rng, n=N, amp=0.05):
    f = np.zeros((5,) + (n[2], n[1], n[0]))
    f[:3] = amp * rng.standard_normal(f[:3].shape)
    f[3] = 0.1 * rng.standard_normal(f[3].shape)
    f[4] = 0.1 * rng.standard_normal(f[4].shape)
    return f


@pytest.mark.parametrize("theta", [0.0, 60.0])
def test_coriolis_is_minus_two_omega_cross_u(theta):
    om = 0.37
    n = (8, 8, 8)
    p = O.make_params(n, lentropy=True, gamma=GAMMA, nu=0.0, omega=om, theta=theta)
    p0 = O.make_params(n, lentropy=True, gamma=GAMMA, nu=0.0)
    f = np.zeros((5, 8, 8, 8))
    u = np.array([0.3, -0.7, 0.45])
    f[:3] = u[:, None, None, None]
    r = O.rhs(f, p)[0] - O.rhs(f, p0)[0]
    th = math.radians(theta)
    Om = om * np.array([-math.sin(th), 0.0, math.cos(th)])
    expect = -2 * np.cross(Om, u)
    for j in range(3):
        np.testing.assert_allclose(r[j], expect[j], rtol=0, atol=1e-15)
    assert abs(np.dot(expect, u)) < 1e-15


def test_uniform_k_profile_equals_chi_const():
    rng = np.random.default_rng(11)
    n = (8, 8, 8)
    f = ent_state(rng, n)
    f[3] = 0.35                                       # uniform density rho = e^0.35
    hc0 = 6e-3
    lay = conv_layers(hcond0=hc0, mpoly0=1.0, mpoly1=1.0, mpoly2=1.0, cool=0.0, cs2cool=0.0)
    pk = O.make_params(n, lentropy=True, gamma=GAMMA, nu=2e-3, layers=lay, dt=0.0)
    pc = O.make_params(n, lentropy=True, gamma=GAMMA, nu=2e-3, chi=hc0 * math.exp(-0.35) / 1.0, dt=0.0)
    rk, dk, _ = O.rhs(f, pk, want_dt=True)
    rc, dc, _ = O.rhs(f, pc, want_dt=True)
    scale = np.abs(rc[4]).max()
    np.testing.assert_allclose(rk[4], rc[4], rtol=0, atol=1e-13 * scale)
    np.testing.assert_allclose(rk[:4], rc[:4], rtol=0, atol=1e-15)
    # the conduction term dominates the Courant step here: nu*dxyz_2 < chi*gamma*dxyz_2
    assert dk == pytest.approx(dc, rel=1e-14)


def kprof_numpy(z, hc0, h1, h2, z1, z2, w):
    return hc0 * (1 + (h1 - 1) * 0.5 * (1 + np.tanh((z - z1) / -w)) + (h2 - 1) * 0.5 * (1 + np.tanh((z - z2) / w)))


def test_k_profile_on_linear_temperature():
    lay = conv_layers(cool=0.0, cs2cool=0.0, widthss=0.15)
    bcz = ["s", "s", "a", "a2", "a2"]
    p = O.make_params(N, lentropy=True, gamma=GAMMA, nu=0.0, L=L, bcz=bcz, layers=lay, z0=Z0)
    dz, z = zgrid()
    g_ss = 0.8
    f = np.zeros((5, N[2], N[1], N[0]))
    f[3] = 0.2
    f[4] = (0.1 + g_ss * (z - Z0))[:, None, None] * np.ones((N[1], N[0]))
    r = O.rhs(f, p)[0]
    cv = 1.0 / GAMMA
    g = g_ss / cv                                     # d lnT/dz
    h = 1e-6
    args = (lay["hcond0"], lay["hcond1"], lay["hcond2"], lay["z1"], lay["z2"], lay["widthss"])
    K = kprof_numpy(z, *args)
    dlnK = (np.log(kprof_numpy(z + h, *args)) - np.log(kprof_numpy(z - h, *args))) / (2 * h)
    expect = math.exp(-0.2) * K * g * (g + dlnK)
    got = r[4][:, 0, 0]
    np.testing.assert_allclose(got, expect, rtol=2e-7, atol=1e-12)
    assert np.abs(r[4] - got[:, None, None]).max() < 1e-13   # horizontally uniform
    # the profile table itself
    prof = O.layer_profile(p)
    np.testing.assert_allclose(prof[:, 0], K, rtol=1e-14)
    np.testing.assert_allclose(prof[:, 1], dlnK, rtol=1e-6, atol=1e-9)


def test_cooling_layer_source():
    lay = conv_layers(hcond0=0.0, cool=15.0, wcool=0.2, cs2cool=0.333333)
    lay["lheatc_kprof"] = 0
    bcz = ["s", "s", "a", "a2", "a2"]
    p = O.make_params(N, lentropy=True, gamma=GAMMA, nu=0.0, L=L, bcz=bcz, layers=lay, z0=Z0)
    dz, z = zgrid()
    f = np.zeros((5, N[2], N[1], N[0]))
    f[3] = 0.15
    f[4] = 0.3                                        # uniform, hot: cs2 > cs2cool
    r = O.rhs(f, p)[0][4][:, 0, 0]
    cp, gm1 = 1.0, GAMMA - 1
    e = O.eos_constants(p)
    cs2 = e["cs20"] * math.exp(e["cv1"] * 0.3 + gm1 * (0.15 - e["lnrho0"]))
    TT = cs2 / (gm1 * cp)
    prof = 15.0 * np.exp(-0.5 * (((Z0 + L[2]) - z) / 0.2) ** 2)
    expect = -(1 / (TT * math.exp(0.15))) * prof * (cs2 - 0.333333) / 0.333333
    assert cs2 > 0.333333 and np.all(r <= 0)
    np.testing.assert_allclose(r, expect, rtol=1e-13, atol=1e-300)


def test_bottom_heating_integrates_to_luminosity():
    nz, Lz = 801, 2.0
    lay = conv_layers(hcond0=0.0, cool=0.0, cs2cool=0.0, luminosity=0.7, wheat=0.1)
    lay["lheatc_kprof"] = 0
    p = O.make_params((8, 6, nz), lentropy=True, gamma=GAMMA, L=(1.3, 0.9, Lz),
                      bcz=["s", "s", "a", "a2", "a2"], layers=lay, z0=-1.0)
    prof = O.layer_profile(p)[:, 2]
    dz = Lz / (nz - 1)
    integral = (prof.sum() - 0.5 * (prof[0] + prof[-1])) * dz * 1.3 * 0.9
    assert integral == pytest.approx(0.7, rel=1e-6)


@pytest.mark.parametrize("side", ["bot", "top"])
def test_c1_ghosts_continue_a_state_with_the_boundary_flux(side):
    lay = conv_layers()
    bcz = ["s", "s", "a", "a2", "c1" if side == "bot" else "a2:c1"]
    if side == "top":
        lay["ftop"] = 0.013
    p = O.make_params(N, lentropy=True, gamma=GAMMA, L=L, bcz=bcz, layers=lay, z0=Z0, gravz=-1.0)
    e = O.eos_constants(p)
    cs20, gm1, cv, lnrho0 = e["cs20"], e["gamma_m1"], e["cv"], e["lnrho0"]
    cp = 1.0
    dz, z = zgrid()
    nb = 0 if side == "bot" else N[2] - 1
    K = lay["hcond0"] * (lay["hcond1"] if side == "bot" else lay["hcond2"])
    F = lay["fbot"] if side == "bot" else lay["ftop"]
    l0, lp, s0 = 0.4, -1.1, 0.05
    lnrho_b = l0 + lp * (z[nb] - Z0)
    # choose ss(boundary) and the slope so that -K dT/dz = F there
    cs2_b = cs20 * math.exp(gm1 * (lnrho_b - lnrho0) + s0 / cv)
    T_b = cs2_b / (gm1 * cp)
    dlnT = -F / (K * T_b)
    sp = cv * dlnT - (cp - cv) * lp                   # cv dlnT = ds + (cp-cv) dlnrho
    f = np.zeros((5, N[2], N[1], N[0]))
    f[3] = (l0 + lp * (z - Z0))[:, None, None]
    f[4] = (s0 + sp * (z - z[nb]))[:, None, None]
    g = O.with_ghosts(f)
    O.boundconds(g, p)
    for i in (1, 2, 3):
        zi = z[nb] - i * dz if side == "bot" else z[nb] + i * dz
        iz = 3 + nb - i if side == "bot" else 3 + nb + i
        np.testing.assert_allclose(g[4, iz], s0 + sp * (zi - z[nb]), rtol=0, atol=1e-13)
        np.testing.assert_allclose(g[3, iz], l0 + lp * (zi - Z0), rtol=0, atol=1e-13)


def test_ct_fixes_the_boundary_temperature():
    lay = conv_layers()
    bcz = ["s", "s", "a", "a2", "c1:cT"]
    p = O.make_params(N, lentropy=True, gamma=GAMMA, L=L, bcz=bcz, layers=lay, z0=Z0, gravz=-1.0)
    e = O.eos_constants(p)
    cs20, gm1, cv1, lnrho0 = e["cs20"], e["gamma_m1"], e["cv1"], e["lnrho0"]
    rng = np.random.default_rng(5)
    f = ent_state(rng)
    g = O.with_ghosts(f)
    O.boundconds(g, p)

    def cs2(iz):
        return cs20 * np.exp(gm1 * (g[3, iz] - lnrho0) + cv1 * g[4, iz])

    top = 3 + N[2] - 1
    np.testing.assert_allclose(cs2(top), p.cs2top, rtol=1e-14)
    for i in (1, 2, 3):
        np.testing.assert_allclose(cs2(top + i) * cs2(top - i), p.cs2top ** 2, rtol=1e-13)
    # the interior is untouched and the bottom plane keeps its value ('c1')
    np.testing.assert_array_equal(g[4, 3:top, 3:-3, 3:-3], f[4, :-1])


def test_conv_slab_step_is_finite_and_uses_every_term():
    """One Courant step of the conv-slab configuration: every added term
    changes the result (so none is silently off)."""
    rng = np.random.default_rng(2)
    f = ent_state(rng, amp=0.02)
    bcz = ["s", "s", "a", "a2", "c1:cT"]
    kw = dict(lentropy=True, gamma=GAMMA, nu=4e-3, L=L, bcz=bcz, z0=Z0, gravz=-1.0, cdt=0.4, cdtv=0.3,
              lupw_lnrho=True, lupw_ss=True, dt=0.0, cs0=0.57735)
    full = O.make_params(N, omega=0.1, layers=conv_layers(), **kw)
    out, t, dt = O.advance(f, full, 1)[:3]
    assert np.all(np.isfinite(out)) and dt > 0
    for change in ({"omega": 0.0}, {"cool": 0.0}, {"hcond0": 0.0}):
        lay = conv_layers()
        om = 0.1
        if "omega" in change:
            om = 0.0
        elif "cool" in change:
            lay["cool"] = 0.0
        else:
            lay["hcond0"] = 0.0
        other = O.advance(f, O.make_params(N, omega=om, layers=lay, **kw), 1)[0]
        assert np.abs(other - out).max() > 1e-9, change
