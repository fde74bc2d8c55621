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

import math

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


def _shell_modes(nx, ny, nz, kmin, kmax):
    """Half-spectrum wavenumbers (kz, ky, kx) with kmin <= |k| <= kmax, in
    the order np.nonzero visits them on the (nz, ny, nx//2+1) grid of FFT
    wavenumbers (so that random draws land on the same modes as in
    bandlimited)."""
    def axis(n):  # wavenumbers with |k| <= kmax in ascending FFT index order
        return list(range(0, kmax + 1)) + list(range(-kmax, 0))
    modes = []
    for kz in axis(nz):
        for ky in axis(ny):
            for kx in range(0, kmax + 1):
                kk = math.sqrt(float(kx * kx) + float(ky * ky) + float(kz * kz))
                if kmin <= kk <= kmax:
                    modes.append((kz, ky, kx))
    return modes


def bandlimited_into(out, kmin, kmax, rms, seed, zslab=64, progress=None):
    """The field of bandlimited(out.shape[0], (nx, ny, nz), kmin, kmax, rms,
    seed), written into out (ncomp, nz, ny, nx; a view is fine) without a
    full-size FFT: with only |k| <= kmax present the inverse transform is a
    sum over few modes, done separably (x, then y, then z slabs as two real
    matrix products), so memory stays at a few slabs.  Needs every extent
    > 2*kmax."""
    ncomp, nz, ny, nx = out.shape
    if min(nx, ny, nz) <= 2 * kmax:
        raise ValueError("grid too small for the shell")
    rng = np.random.default_rng(seed)
    modes = _shell_modes(nx, ny, nz, kmin, kmax)
    m = len(modes)
    ks = list(range(-kmax, kmax + 1))        # z and y wavenumbers, array index k + kmax
    two_pi = 2.0 * math.pi
    fx = np.exp(1j * two_pi * np.outer(np.arange(kmax + 1), np.arange(nx)) / nx)     # (kx, x)
    fy = np.exp(1j * two_pi * np.outer(ks, np.arange(ny)) / ny)                       # (ky, y)
    fz = np.exp(1j * two_pi * np.outer(ks, np.arange(nz)) / nz)                       # (kz, z)
    for c in range(ncomp):
        coef = rng.standard_normal(m) + 1j * rng.standard_normal(m)
        cm = np.zeros((len(ks), len(ks), kmax + 1), dtype=np.complex128)
        for (kz, ky, kx), v in zip(modes, coef):
            # irfft counts kx > 0 twice (conjugate half) and the real part of kx = 0 once
            cm[kz + kmax, ky + kmax, kx] = v * (2.0 if kx > 0 else 1.0)
        g = cm @ fx                                          # (kz, ky, x)
        h = np.einsum("jy,ijx->iyx", fy, g).reshape(len(ks), ny * nx)
        # Re(sum_kz e^{i kz z} h_kz) = [cos | -sin](z, kz) @ [Re h ; Im h]
        hs = np.concatenate([h.real, h.imag])                # (2 kz, y x)
        del g, h
        buf = np.empty((zslab, ny * nx))
        for z0 in range(0, nz, zslab):
            z1 = min(nz, z0 + zslab)
            a = np.concatenate([fz[:, z0:z1].real, -fz[:, z0:z1].imag]).T
            np.matmul(a, hs, out=buf[: z1 - z0])
            out[c, z0:z1] = buf[: z1 - z0].reshape(z1 - z0, ny, nx)
        if progress:
            progress(f"band-limited component {c + 1}/{ncomp} done")
    ssq = 0.0
    for c in range(ncomp):
        for z0 in range(0, nz, zslab):
            blk = out[c, z0:min(nz, z0 + zslab)]
            ssq += float(np.einsum("zyx,zyx->", blk, blk))
    cur = math.sqrt(ssq / (float(nx) * ny * nz))
    if cur > 0:
        for c in range(ncomp):
            for z0 in range(0, nz, zslab):
                out[c, z0:min(nz, z0 + zslab)] *= rms / cur


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


def _step(x, x0, w):
    """step (sub.f90:4975-4987), tini = 5*tiny."""
    return 0.5 * (1 + np.tanh((x - x0) / (w + 5 * 2.2250738585072014e-308)))


def _polytrope_layer(f, z, kind, mpoly, zint, zbot, zblend, isoth, cs2int, vint, gamma, gravz, width,
                     cp=1.0, fac_cs=None):
    """One layer of polytropic_lnrho_z (density.f90:2111-2177, kind 'lnrho')
    or polytropic_ss_z (entropy.f90:2186-2260, kind 'ss'), called from the
    top layer down; f is the (nz,) profile, blended in below zblend with
    step(z, zblend, width).  Returns the new profile, cs2int and the value at
    the layer bottom."""
    epsi = 5 * np.finfo(np.float64).eps
    gamma_m1, cp1 = gamma - 1.0, 1.0 / cp
    stp = _step(z, zblend, width)
    if isoth:
        beta1 = 0.0
        if kind == "lnrho":
            tmp = vint + gamma * gravz / cs2int * (z - zint)
        else:
            tmp = vint - gamma_m1 * gravz * (z - zint) / cs2int / cp1
    else:
        beta1 = gamma * gravz / (mpoly + 1)
        t = np.maximum(1.0 + beta1 * (z - zint) / cs2int, epsi)
        tmp = vint + mpoly * np.log(t) if kind == "lnrho" else vint + (1 - mpoly * gamma_m1) / gamma * np.log(t) / cp1
    f = stp * f + (1 - stp) * tmp
    if isoth:
        if kind == "lnrho":
            vint = vint + gamma * gravz / cs2int * (zbot - zint)
        else:
            vint = vint - gamma_m1 * gravz * (zbot - zint) / cs2int / cp1
            if fac_cs is not None:
                vint = vint + 2 * np.log(fac_cs) / gamma / cp1
    elif kind == "lnrho":
        vint = vint + mpoly * np.log(1 + beta1 * (zbot - zint) / cs2int)
    else:
        vint = vint + (1 - mpoly * gamma_m1) / gamma * np.log(1 + beta1 * (zbot - zint) / cs2int) / cp1
    cs2int = fac_cs ** 2 * cs2int if (isoth and fac_cs is not None) else cs2int + beta1 * (zbot - zint)
    return f, cs2int, vint


def conv_slab_state(n=32, nz=None, cp=1.0):
    """The initial state of samples/conv-slab (start.in) on an n x n x nz grid
    (default nz = n): the 'piecew-poly' stratification of lnrho (density.f90:
    1578-1605) and ss (entropy.f90:1971-1993) -- an isothermal top layer
    (isothtop = 1) above z2 = 1, a polytrope of index mpoly0 = 1 between
    z1 = 0 and z2, mpoly1 = 3 below, referred to zref = 1.32 where cs = cs0 =
    0.57735, rho = 1, ss = 0, blended with width 0.05 -- and the 'up-down'
    u_z (hydro.f90:2633-2647, ampluu = 0.1, widthuu = 0.1, about z1 = 0).
    Box xyz0 = (-0.5, -0.5, -0.68), Lxyz = (1, 1, 2), periodic in x and y
    (cell centres), z points from -0.68 to 1.32.  cp scales ss (the
    samples/conv-slab_cp_2 variant runs cp = 2).  Returns (5, nz, n, n)."""
    nz = n if nz is None else nz
    gamma, gravz, cs0 = 1.6666666, -1.0, 0.57735
    x0, y0, z0, Lx, Ly, Lz = -0.5, -0.5, -0.68, 1.0, 1.0, 2.0
    zref, z1, z2, ztop = 1.32, 0.0, 1.0, z0 + Lz
    mpoly0, mpoly1, mpoly2 = 1.0, 3.0, 0.0
    x = x0 + (np.arange(n) + 0.5) * (Lx / n)
    y = y0 + (np.arange(n) + 0.5) * (Ly / n)
    z = z0 + np.arange(nz) * (Lz / (nz - 1))
    # lnrho: top (isothermal), unstable, stable layer
    lnrho = np.zeros(nz)
    cs2int, lint = cs0 ** 2, 0.0
    lnrho, cs2int, lint = _polytrope_layer(lnrho, z, "lnrho", mpoly2, zref, z2, ztop + Lz, 1, cs2int, lint,
                                           gamma, gravz, 0.05, fac_cs=1.0)
    lnrho, cs2int, lint = _polytrope_layer(lnrho, z, "lnrho", mpoly0, z2, z1, z2, 0, cs2int, lint, gamma, gravz, 0.05)
    lnrho, cs2int, lint = _polytrope_layer(lnrho, z, "lnrho", mpoly1, z1, z0, z1, 0, cs2int, lint, gamma, gravz, 0.05)
    ss = np.zeros(nz)
    cs2int, sint = cs0 ** 2, 0.0
    ss, cs2int, sint = _polytrope_layer(ss, z, "ss", mpoly2, zref, z2, z0 + 2 * Lz, 1, cs2int, sint, gamma, gravz,
                                        0.05, cp=cp, fac_cs=1.0)
    ss, cs2int, sint = _polytrope_layer(ss, z, "ss", mpoly0, z2, z1, z2, 0, cs2int, sint, gamma, gravz, 0.05, cp=cp)
    ss, cs2int, sint = _polytrope_layer(ss, z, "ss", mpoly1, z1, z0, z1, 0, cs2int, sint, gamma, gravz, 0.05, cp=cp)
    f = np.zeros((5, nz, n, n))
    f[3] = lnrho[:, None, None]
    f[4] = ss[:, None, None]
    w = 0.1
    prof = 0.1 * np.exp(-0.5 * (z - z1) ** 2 / w ** 2)[:, None, None]
    X, Y = x[None, None, :], y[None, :, None]
    t1 = np.sqrt((X - (x0 + 0.3 * Lx)) ** 2 + (Y - (y0 + 0.3 * Ly)) ** 2)
    t2 = np.sqrt((X - (x0 + 0.5 * Lx)) ** 2 + (Y - (y0 + 0.8 * Ly)) ** 2)
    f[2] = prof * np.exp(-0.5 * (t1 ** 2) / w ** 2)
    f[2] = f[2] - 0.7 * prof * np.exp(-0.5 * (t2 ** 2) / w ** 2)
    return f


def fluxrings_state(n=32):
    """The initial state of samples/interlocked-fluxrings (start.in): two
    interlocked magnetic flux rings in A (initial_condition/fluxrings.f90:
    fluxrings :204-290, norm_ring 'tanh' profile :300-360), flux 1, radius
    1, width 0.3, ring 1 about the z axis displaced to (0, -0.5, 0), ring 2
    about the x axis displaced to (0, 0.5, 0), no current (Iring = 0); u = 0,
    lnrho = 0; periodic box [-pi, pi]^3 (cell centres).  Returns (7, n, n, n)
    in the isothermal MHD slots uu, lnrho, aa."""
    epsi = 5 * np.finfo(np.float64).eps
    x = -np.pi + (np.arange(n) + 0.5) * (2 * np.pi / n)
    X, Y, Z = x[None, None, :], x[None, :, None], x[:, None, None]
    f = np.zeros((7, n, n, n))
    for axis, disp in (((0.0, 0.0, 1.0), (0.0, -0.5, 0.0)), ((1.0, 0.0, 0.0), (0.0, 0.5, 0.0))):
        fring, R0, width, ampl = 1.0, 1.0, 0.3, 1.0
        phi = np.arctan2(axis[1], axis[0] + epsi)
        theta = np.arctan2(np.sqrt(axis[0] ** 2 + axis[1] ** 2) + epsi, axis[2])
        ct, st, cp, sp = np.cos(theta), np.sin(theta), np.cos(phi), np.sin(phi)
        dx, dy, dz = X - disp[0], Y - disp[1], Z - disp[2]
        xx1 = ct * cp * dx + ct * sp * dy - st * dz
        yy1 = -sp * dx + cp * dy
        zz1 = st * cp * dx + st * sp * dy + ct * dz
        tmp = np.sqrt(xx1 ** 2 + yy1 ** 2) - R0
        v3 = -(fring * 0.5 * (1 + np.tanh(tmp / width)) * 0.5 / width / np.cosh(zz1 / width) ** 2)
        tmp = 0.0 * 0.5 * (1 + np.tanh((width - np.sqrt(tmp ** 2 + zz1 ** 2)) / width))   # Iring = 0
        ph = np.arctan2(yy1, xx1)
        v1, v2 = -tmp * np.sin(ph), tmp * np.cos(ph)
        f[4] = f[4] + ampl * (+ct * cp * v1 - sp * v2 + st * cp * v3)
        f[5] = f[5] + ampl * (+ct * sp * v1 + cp * v2 + st * sp * v3)
        f[6] = f[6] + ampl * (-st * v1 + ct * v3)
    return f
