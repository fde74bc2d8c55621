"""Pins the oracle's upwinding and helical forcing against what the
mathematics fixes.

* Upwinding (lupw_uu / lupw_lnrho: u_dot_grad with upwind=T,
  sub.f90:7541-7574, der6 with UPWIND, deriv.f90): the correction subtracts
  sum_j |u_j| D6_j f from u.grad f, where D6_j is the 7-point sixth
  difference divided by 60 h_j.  On a Fourier mode its symbol is
  -(64 / (60 h_j)) sin^6(k_j h_j / 2), so the change of the RHS is a damping
  -sum_j |u_j| (64 / (60 h_j)) sin^6(k_j h_j / 2) f, pointwise: per
  direction, with the speed of that direction, on unequal spacings.
* Helical forcing (forcing_hel / fconst_coefs_hel, forcing.f90:1773-2474):
  with relhel = +-1 the force is a Beltrami field, curl F = +-k1_ff |k| F,
  div F = 0, of uniform magnitude force*sqrt(dt kav cs0eff^3 / 2); a phase is
  a translation along x.  Checked with spectral derivatives of the oracle's
  addforce output."""
import math

import numpy as np
import pytest

import oracle as O
import pcmx_forcing as PF

U0 = (0.3, -0.2, 0.15)
NS = (16, 12, 20)
LS = (2 * math.pi, 3.0, 2.5)


def coords(n, L):
    hs = [L[a] / n[a] for a in range(3)]
    x = hs[0] * np.arange(n[0])
    y = hs[1] * np.arange(n[1])
    z = hs[2] * np.arange(n[2])
    return hs, x[None, None, :], y[None, :, None], z[:, None, None]


def d6_symbol(k, h):
    return -(64.0 / (60.0 * h)) * math.sin(k * h / 2) ** 6


def mode(m):
    hs, X, Y, Z = coords(NS, LS)
    ks = [2 * math.pi * m[a] / LS[a] for a in range(3)]
    return hs, ks, np.sin(ks[0] * X + ks[1] * Y + ks[2] * Z)


@pytest.mark.parametrize("m", [(1, 2, 3), (3, 0, 1), (0, 5, 0)])
def test_upwinded_continuity_is_a_damping_per_direction(m):
    hs, ks, s = mode(m)
    A = 0.01
    f = np.zeros((4,) + s.shape)
    for j in range(3):
        f[j] = U0[j]
    f[3] = A * s
    off = O.rhs(f, O.make_params(NS, L=LS))[0]
    on = O.rhs(f, O.make_params(NS, L=LS, lupw_lnrho=True))[0]
    damp = sum(abs(U0[j]) * d6_symbol(ks[j], hs[j]) for j in range(3))
    np.testing.assert_allclose(on[3] - off[3], damp * A * s, rtol=0, atol=1e-15)
    np.testing.assert_array_equal(on[:3], off[:3])


@pytest.mark.parametrize("i", [0, 1, 2])
def test_upwinded_advection_uses_the_speed_of_each_direction(i):
    # u = U0 + eps e_i sin(k.x): |u_j| is constant for j != i and varies for j = i
    hs, ks, s = mode((1, 2, 3))
    eps = 0.01
    f = np.zeros((4,) + s.shape)
    for j in range(3):
        f[j] = U0[j] + (eps * s if j == i else 0.0)
    off = O.rhs(f, O.make_params(NS, L=LS))[0]
    on = O.rhs(f, O.make_params(NS, L=LS, lupw_uu=True))[0]
    speed = [abs(U0[j] + (eps * s if j == i else 0.0)) for j in range(3)]
    expect = sum(speed[j] * d6_symbol(ks[j], hs[j]) for j in range(3)) * eps * s
    np.testing.assert_allclose(on[i] - off[i], expect, rtol=0, atol=1e-15)
    for c in range(3):
        if c != i:
            assert np.abs(on[c] - off[c]).max() < 1e-15   # D6 of a constant
    np.testing.assert_array_equal(on[3], off[3])


def test_upwinding_damps_a_step_and_vanishes_for_resolved_fields():
    # a resolved mode is barely touched (sin^6(kh/2) ~ (kh/2)^6), a grid-scale one is damped
    hs, ks, s = mode((1, 1, 1))
    f = np.zeros((4,) + s.shape)
    f[0] = 0.2
    f[3] = 1e-3 * s
    p_on = O.make_params(NS, L=LS, lupw_lnrho=True, dt=1e-3)
    p_off = O.make_params(NS, L=LS, dt=1e-3)
    a = O.advance(f, p_on, 1)[0]
    b = O.advance(f, p_off, 1)[0]
    rel = np.abs(a[3] - b[3]).max() / np.abs(b[3]).max()
    assert 0 < rel < 1e-6
    hs, ks, s = mode((NS[0] // 2 - 1, 0, 0))
    f[3] = 1e-3 * s
    a = O.advance(f, p_on, 1)[0]
    b = O.advance(f, p_off, 1)[0]
    assert np.abs(a[3]).max() < np.abs(b[3]).max()


# ------------------------------------------------------------------ forcing
def curl_div(F, L):
    n = F.shape[1:]
    kz = 2 * np.pi * np.fft.fftfreq(n[0], L[2] / n[0])[:, None, None]
    ky = 2 * np.pi * np.fft.fftfreq(n[1], L[1] / n[1])[None, :, None]
    kx = 2 * np.pi * np.fft.fftfreq(n[2], L[0] / n[2])[None, None, :]
    Fh = [np.fft.fftn(F[j]) for j in range(3)]
    d = lambda k, g: np.real(np.fft.ifftn(1j * k * g))  # noqa: E731
    curl = np.array([d(ky, Fh[2]) - d(kz, Fh[1]), d(kz, Fh[0]) - d(kx, Fh[2]), d(kx, Fh[1]) - d(ky, Fh[0])])
    div = d(kx, Fh[0]) + d(ky, Fh[1]) + d(kz, Fh[2])
    return curl, div


def forced_zero_state(n, kvec, relhel, k1_ff=1.0, phase=0.7, phi=1.1, force=0.07, dt=2e-3, kav=3.0,
                      xyz0=(-math.pi,) * 3, target=0, lmagnetic=False):
    p = O.make_params(n, lmagnetic=lmagnetic)
    pars = PF.helical_pars(kvec, phase, phi, force, dt, kav, relhel=relhel)
    fo = O.make_forcing(pars["kk"], pars["phase"], pars["fact"], pars["coef1"], pars["coef2"], pars["fda"],
                        k1_ff=k1_ff, xyz0=xyz0, target=target)
    out = O.addforce(np.zeros((O.nvar(p), n, n, n)), p, fo)
    return out, pars


@pytest.mark.parametrize("kvec,k1_ff", [((1, 2, -2), 1.0), ((3, 0, 0), 1.0), ((0, -1, 1), 2.0), ((2, 1, 0), 1.0)])
@pytest.mark.parametrize("relhel", [1.0, -1.0])
def test_helical_force_is_beltrami(kvec, k1_ff, relhel):
    n = 24
    force, dt, kav = 0.07, 2e-3, 3.0
    out, pars = forced_zero_state(n, kvec, relhel, k1_ff=k1_ff, force=force, dt=dt, kav=kav)
    F = out[0:3]
    assert np.abs(out[3]).max() == 0.0
    kl = k1_ff * math.sqrt(sum(v * v for v in kvec))
    curl, div = curl_div(F, (2 * math.pi,) * 3)
    scale = np.abs(F).max()
    np.testing.assert_allclose(curl, relhel * kl * F, rtol=0, atol=1e-12 * kl * scale)
    assert np.abs(div).max() < 1e-12 * kl * scale
    mag = np.sqrt((F ** 2).sum(axis=0))
    np.testing.assert_allclose(mag, force * math.sqrt(dt * kav / 2), rtol=1e-13)


def test_helical_coefficients_geometry():
    for kvec in [(1, 2, -2), (4, 0, 0), (0, 0, 3), (1, 1, 1), (-2, 3, 5)]:
        for phi in (0.0, 0.4, 2.9, 5.5):
            p = PF.helical_pars(kvec, 0.3, phi, 0.05, 1e-3, 4.0, relhel=0.5)
            k = np.array(kvec, float)
            c1, c2 = np.array(p["coef1"]), np.array(p["coef2"])
            tol = 1e-13 * (k @ k) ** 1.5
            assert abs(c1 @ k) < tol and abs(c2 @ k) < tol and abs(c1 @ c2) < tol * (k @ k) ** 0.5
            np.testing.assert_allclose(np.linalg.norm(c2), 0.5 * np.linalg.norm(c1), rtol=1e-13)
            np.testing.assert_allclose(np.linalg.norm(c1), (k @ k), rtol=1e-13)   # k |k x e|, |e| = 1, e.k = 0
            # (coef1 + i coef2) is an eigenvector of i k x . with eigenvalue |k| (relhel = 1 part)
    q = PF.helical_pars((1, 2, -2), 0.3, 0.4, 0.05, 1e-3, 4.0, scale_sqrt_dt=True)
    r = PF.helical_pars((1, 2, -2), 0.3, 0.4, 0.05, 1e-3, 4.0)
    assert q["fact"] * math.sqrt(1e-3) == r["fact"]


def test_phase_is_a_translation_along_x():
    n, kvec = 16, (2, 1, -1)
    a, pars = forced_zero_state(n, kvec, 1.0, phase=0.0)
    shift = 2 * (2 * math.pi / n)                                   # two cells
    b, _ = forced_zero_state(n, kvec, 1.0, phase=kvec[0] * shift)   # exp(i k1 (x + shift))
    np.testing.assert_allclose(b, np.roll(a, -2, axis=3), rtol=0, atol=1e-15)


def test_forcing_targets_the_vector_potential():
    n = 12
    p = O.make_params(n, lmagnetic=True)
    out, _ = forced_zero_state(n, (1, 0, 1), 1.0, target=O.iax(p), lmagnetic=True)
    assert np.abs(out[:O.iax(p)]).max() == 0.0
    assert np.abs(out[O.iax(p):O.iax(p) + 3]).max() > 0.0


def test_k_dat_reader_and_shell(tmp_path):
    kav, vecs = PF.shell_vectors(1.0, 2.0)
    path = tmp_path / "k.dat"
    with open(path, "w") as fh:
        fh.write(f"{len(vecs)} {kav}\n")
        for c in range(3):
            fh.write(" ".join(f"{v:.0f}" for v in vecs[:, c]) + "\n")
    kav2, vecs2 = PF.read_k_dat(str(path))
    assert len(vecs) == 26                                          # 6 of length 1, 12 of sqrt 2, 8 of sqrt 3
    assert abs(kav2 - (6 + 12 * math.sqrt(2) + 8 * math.sqrt(3)) / 26) < 1e-15
    np.testing.assert_array_equal(vecs2, vecs)
    rng = np.random.default_rng(0)
    for _ in range(100):
        ik, phase, phi = PF.draw(rng, len(vecs))
        assert 0 <= ik < len(vecs) and -math.pi <= phase < math.pi and 0 <= phi < 2 * math.pi


@pytest.mark.parametrize("m", [(1, 2, 3), (0, 4, 1)])
def test_upwinded_entropy_advection_is_a_damping_per_direction(m):
    """lupw_ss (entropy.f90:3576): the same damping on ss, nothing else changes."""
    hs, ks, s = mode(m)
    A = 0.01
    f = np.zeros((5,) + s.shape)
    for j in range(3):
        f[j] = U0[j]
    f[4] = A * s
    kw = dict(L=LS, lentropy=True, gamma=5 / 3, chi=1e-3)
    off = O.rhs(f, O.make_params(NS, **kw))[0]
    on = O.rhs(f, O.make_params(NS, lupw_ss=True, **kw))[0]
    damp = sum(abs(U0[j]) * d6_symbol(ks[j], hs[j]) for j in range(3))
    np.testing.assert_allclose(on[4] - off[4], damp * A * s, rtol=0, atol=1e-15)
    np.testing.assert_array_equal(on[:4], off[:4])
