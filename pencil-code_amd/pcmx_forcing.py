"""Host side of helical forcing: the parameters the library's forcing term
takes (pcmx.set_forcing), formed as the reference forms them.

* ``read_k_dat``: the wave-vector table k.dat (forcing.f90:888-899: nk and
  kav, then the kx, ky, kz lists, list-directed).
* ``helical_pars``: fconst_coefs_hel (forcing.f90:1773-2001) on its default
  path (no shear, no k scaling, the new e-vector, force_strength=0) for
  given random draws: the chosen wave vector, the phase and the angle phi
  that the reference draws with its random_number_wrapper.  Returns kk,
  phase, fact, coef1, coef2, fda, the arguments of pcmx.set_forcing.
* ``draw``: draws (ik, phase, phi) from a numpy generator in the reference's
  ranges (phase = pi*(2 r1 - 1), ik = int(nk*0.9999*r2), phi = 2 pi r3).  It
  is not the reference's generator: a Fortran caller passes the reference's
  own draws through forcing_pars_hel instead.
"""
from __future__ import annotations

import math

import numpy as np


def read_k_dat(path):
    """(kav, kvecs[nk, 3]) from a k.dat file."""
    with open(path) as fh:
        tok = fh.read().split()
    nk, kav = int(tok[0]), float(tok[1])
    vals = np.array([float(t) for t in tok[2:2 + 3 * nk]])
    if vals.size != 3 * nk:
        raise ValueError(f"{path}: {vals.size} wave-vector components for nk={nk}")
    return kav, vals.reshape(3, nk).T.copy()


def _cross(a, b):
    return (a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0])


def helical_pars(kvec, phase, phi, force, dt, kav, relhel=1.0, cs0eff=1.0, slope_ff=0.0, scale_sqrt_dt=False):
    """Parameters of one helical forcing step.  kvec: the selected wave
    vector (kkx(ik), kky(ik), kkz(ik)); phase in [-pi, pi); phi in [0, 2 pi).
    With scale_sqrt_dt the returned fact is force/ffnorm, to be multiplied by
    sqrt(dt) of the step by the library; otherwise force/ffnorm*sqrt(dt)."""
    kx, ky, kz = (float(v) for v in kvec)
    k2 = kx ** 2 + ky ** 2 + kz ** 2
    k = math.sqrt(k2)
    ee = (0.0, 1.0, 0.0) if (ky == 0 and kz == 0) else (1.0, 0.0, 0.0)
    kk = (kx, ky, kz)
    e1 = _cross(kk, ee)
    nrm = math.sqrt(e1[0] ** 2 + e1[1] ** 2 + e1[2] ** 2)
    e1 = tuple(v / nrm for v in e1)
    e2 = _cross(kk, e1)
    nrm = math.sqrt(e2[0] ** 2 + e2[1] ** 2 + e2[2] ** 2)
    e2 = tuple(v / nrm for v in e2)
    ex, ey, ez = (math.cos(phi) * a + math.sin(phi) * b for a, b in zip(e1, e2))
    kde = kx * ex + ky * ey + kz * ez
    kex, key, kez = ky * ez - kz * ey, kz * ex - kx * ez, kx * ey - ky * ex
    kkex, kkey, kkez = ky * kez - kz * key, kz * kex - kx * kez, kx * key - ky * kex
    ffnorm = (math.sqrt(1.0 + relhel ** 2) * k * math.sqrt(k2 - kde ** 2) / math.sqrt(kav * cs0eff ** 3)
              * (k / kav) ** slope_ff)
    fact = force / ffnorm if scale_sqrt_dt else force / ffnorm * math.sqrt(dt)
    return dict(kk=kk, phase=float(phase), fact=fact, coef1=(k * kex, k * key, k * kez),
                coef2=(relhel * kkex, relhel * kkey, relhel * kkez), fda=(1.0, 1.0, 1.0),
                scale_sqrt_dt=bool(scale_sqrt_dt))


def draw(rng, nk):
    """(ik 0-based, phase, phi) in the reference's ranges."""
    r1, r2 = rng.random(), rng.random()
    phase = math.pi * (2 * r1 - 1.0)
    ik = int(nk * (0.9999 * r2))
    phi = rng.random() * 2 * math.pi
    return ik, phase, phi


def shell_vectors(kmin, kmax):
    """Integer wave vectors with kmin <= |k| < kmax and their mean length,
    the content of a k.dat made for a periodic 2 pi box."""
    r = int(math.ceil(kmax))
    g = np.arange(-r, r + 1)
    kx, ky, kz = np.meshgrid(g, g, g, indexing="ij")
    kl = np.sqrt(kx ** 2 + ky ** 2 + kz ** 2)
    sel = (kl >= kmin) & (kl < kmax)
    vecs = np.stack([kx[sel], ky[sel], kz[sel]], axis=1).astype(float)
    return float(kl[sel].mean()), vecs
