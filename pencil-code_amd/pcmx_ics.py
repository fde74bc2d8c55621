"""Seeded synthetic initial conditions for the benchmark and parity cases.

* ``gaussian_noise``: i.i.d. normal values of a given amplitude per cell and
  component, the shape of the reference's ``inituu='gaussian-noise'``
  (initcond.f90:5005-5047).  The draw comes from numpy's PCG64 with the given
  seed, not from the reference's mars_ran generator.
* ``bandlimited``: a random vector field whose Fourier modes are confined to
  the shell kmin <= |k| <= kmax (integer wavenumbers of the 2pi box), scaled
  to a given rms, as the isothermal-MHD cases ask for (the box-mean mode is
  outside every shell, so the field has zero mean).
"""
from __future__ import annotations

import numpy as np


def gaussian_noise(ncomp, n, ampl, seed):
    nx, ny, nz = (n, n, n) if np.isscalar(n) else n
    rng = np.random.default_rng(seed)
    return ampl * rng.standard_normal((ncomp, nz, ny, nx))


def bandlimited(ncomp, n, kmin, kmax, rms, seed):
    """Random field with Fourier modes only on the shell kmin <= |k| <= kmax:
    i.i.d. complex normal coefficients on the shell modes of the half
    spectrum, one inverse real FFT per component (the same statistics as
    shell-filtered white noise, without the forward transform), scaled so
    that sqrt(<|f|^2>) = rms."""
    nx, ny, nz = (n, n, n) if np.isscalar(n) else n
    rng = np.random.default_rng(seed)
    kx = np.arange(nx // 2 + 1)[None, None, :]
    ky = np.fft.fftfreq(ny, 1.0 / ny)[None, :, None]
    kz = np.fft.fftfreq(nz, 1.0 / nz)[:, None, None]
    kk = np.sqrt(kx * kx + ky * ky + kz * kz)
    idx = np.nonzero((kk >= kmin) & (kk <= kmax))
    out = np.empty((ncomp, nz, ny, nx))
    spec = np.zeros((nz, ny, nx // 2 + 1), dtype=np.complex128)
    for c in range(ncomp):
        m = idx[0].size
        spec[idx] = rng.standard_normal(m) + 1j * rng.standard_normal(m)
        out[c] = np.fft.irfftn(spec, s=(nz, ny, nx), axes=(0, 1, 2))
        spec[idx] = 0
    cur = np.sqrt(np.mean(np.sum(out * out, axis=0)))
    return out * (rms / cur) if cur > 0 else out


def config_state(kind, n, seed, lmagnetic=False, lentropy=False):
    """Interior state (mvar, nz, ny, nx) in the registration order
    uu, lnrho, [ss], [aa].

    kind 'noise': u ~ N(0, 1e-3^2) per component, lnrho = 0, ss ~ N(0, 1e-3^2),
                  A ~ N(0, 1e-3^2) (the hydro / entropy / large-box recipe);
    kind 'shells': u and A band-limited to 1 <= |k| <= 4 with rms 0.05 and
                  0.02, lnrho = 0 (the isothermal-MHD recipe)."""
    nv = 4 + (1 if lentropy else 0) + (3 if lmagnetic else 0)
    nx, ny, nz = (n, n, n) if np.isscalar(n) else n
    f = np.zeros((nv, nz, ny, nx))
    ia = (5 if lentropy else 4)
    if kind == "noise":
        f[0:3] = gaussian_noise(3, n, 1e-3, seed)
        if lentropy:
            f[4] = gaussian_noise(1, n, 1e-3, seed + 1)[0]
        if lmagnetic:
            f[ia:ia + 3] = gaussian_noise(3, n, 1e-3, seed + 2)
    elif kind == "shells":
        f[0:3] = bandlimited(3, n, 1, 4, 0.05, seed)
        if lentropy:
            f[4] = gaussian_noise(1, n, 1e-3, seed + 1)[0]
        if lmagnetic:
            f[ia:ia + 3] = bandlimited(3, n, 1, 4, 0.02, seed + 2)
    else:
        raise ValueError(kind)
    return f
