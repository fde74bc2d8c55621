"""ctypes wrapper of the CPU oracle (``oracle/pc_oracle.c``).

TEST INFRASTRUCTURE: only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` import this module, and only as the
checker.  The product (``libpcmx``) never loads it.

Arrays follow the reference f-array: ``f(mx,my,mz,nvar)`` column-major, which
is a C-ordered numpy array of shape ``(nvar, mz, my, mx)``.
"""
from __future__ import annotations

import ctypes
import math
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libpcoracle.so")
NGHOST = 3


class Params(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
        ("lmagnetic", ctypes.c_int), ("lentropy", ctypes.c_int),
        ("Lx", ctypes.c_double), ("Ly", ctypes.c_double), ("Lz", ctypes.c_double),
        ("nu", ctypes.c_double), ("eta", ctypes.c_double), ("chi", ctypes.c_double),
        ("cs0", ctypes.c_double), ("gamma", ctypes.c_double), ("cp", ctypes.c_double),
        ("rho0", ctypes.c_double), ("mu0", ctypes.c_double),
        ("cdt", ctypes.c_double), ("cdtv", ctypes.c_double),
        ("dt", ctypes.c_double),
        ("lupw_uu", ctypes.c_int), ("lupw_lnrho", ctypes.c_int),
        ("gravz", ctypes.c_double), ("lperi_z", ctypes.c_int),
        ("bcz_bot", ctypes.c_int * 8), ("bcz_top", ctypes.c_int * 8),
        ("lupw_ss", ctypes.c_int),
        ("omega", ctypes.c_double), ("theta", ctypes.c_double),
        ("lheatc_kprof", ctypes.c_int), ("cooling_profile", ctypes.c_int),
        ("hcond0", ctypes.c_double), ("hcond1", ctypes.c_double), ("hcond2", ctypes.c_double),
        ("widthss", ctypes.c_double), ("z1", ctypes.c_double), ("z2", ctypes.c_double),
        ("luminosity", ctypes.c_double), ("wheat", ctypes.c_double),
        ("cool", ctypes.c_double), ("wcool", ctypes.c_double), ("cs2cool", ctypes.c_double),
        ("fbot", ctypes.c_double), ("ftop", ctypes.c_double),
        ("cs2bot", ctypes.c_double), ("cs2top", ctypes.c_double),
        ("z0", ctypes.c_double),
    ]


BC_CODES = {"s": 1, "a": 2, "a2": 3, "c1": 4, "cT": 5}

# conv-slab layer fields of Params (set through make_params(layers=...))
LAYER_FIELDS = ("lheatc_kprof", "cooling_profile", "hcond0", "hcond1", "hcond2", "widthss", "z1", "z2",
                "luminosity", "wheat", "cool", "wcool", "cs2cool", "fbot", "ftop", "cs2bot", "cs2top")


def layer_defaults(gamma, cp, gravz, bcz_ss, hcond0=0.0, mpoly0=1.5, mpoly1=1.5, mpoly2=1.5, widthss=0.05,
                   z1=0.0, z2=1.0, luminosity=0.0, wheat=0.1, cool=0.0, wcool=0.1, cs2cool=0.0,
                   cooling_profile="gaussian", cs2top=None, cs2bot=None, fbot=None, ftop=None,
                   kprof=True):
    """The conv-slab layer fields as initialize_energy derives them
    (entropy.f90:805-890): hcond1 = (mpoly1+1)/(mpoly0+1), hcond2 likewise;
    Fbot (Ftop) = -gamma/(cp*gamma_m1)*hcond0*gravz/(mpoly0+1) when the bottom
    (top) condition of ss is 'c1', else 0; with cooling, cs2top =
    (cool*cs2cool + cool2*cs2cool2)/(cool+cool2), cool2 = cs2cool2 = 0.
    bcz_ss is the run.in entry of ss, e.g. 'c1:cT'."""
    lo, _, hi = bcz_ss.partition(":")
    hi = hi or lo
    gamma_m1 = gamma - 1.0
    flux = -gamma / (cp * gamma_m1) * hcond0 * gravz / (mpoly0 + 1)
    if cs2top is None:
        cs2top = (cool * cs2cool + 0.0 * 0.0) / (cool + 0.0) if cool != 0.0 else 0.0
    return dict(lheatc_kprof=int(kprof), cooling_profile={"gaussian": 0, "step": 1}[cooling_profile],
                hcond0=hcond0, hcond1=(mpoly1 + 1.) / (mpoly0 + 1.), hcond2=(mpoly2 + 1.) / (mpoly0 + 1.),
                widthss=widthss, z1=z1, z2=z2, luminosity=luminosity, wheat=wheat, cool=cool, wcool=wcool,
                cs2cool=cs2cool, fbot=(flux if lo == "c1" else 0.0) if fbot is None else fbot,
                ftop=(flux if hi == "c1" else 0.0) if ftop is None else ftop,
                cs2bot=0.0 if cs2bot is None else cs2bot, cs2top=cs2top)


def bc_codes(bcz, nv):
    """Per-variable (bottom, top) codes from run.in style entries: 'a2' or
    'a2:s' (bottom:top)."""
    if len(bcz) != nv:
        raise ValueError(f"{len(bcz)} z boundary conditions for {nv} variables")
    bot, top = (ctypes.c_int * 8)(), (ctypes.c_int * 8)()
    for k, b in enumerate(bcz):
        lo, _, hi = b.partition(":")
        bot[k], top[k] = BC_CODES[lo], BC_CODES[hi or lo]
    return bot, top


class Forcing(ctypes.Structure):
    _fields_ = [
        ("kk", ctypes.c_double * 3), ("phase", ctypes.c_double), ("fact", ctypes.c_double),
        ("coef1", ctypes.c_double * 3), ("coef2", ctypes.c_double * 3), ("fda", ctypes.c_double * 3),
        ("k1_ff", ctypes.c_double), ("x0", ctypes.c_double), ("y0", ctypes.c_double), ("z0", ctypes.c_double),
        ("target", ctypes.c_int),
    ]


def make_forcing(kk, phase, fact, coef1, coef2, fda=(1.0, 1.0, 1.0), k1_ff=1.0, xyz0=(-math.pi,) * 3, target=0):
    return Forcing((ctypes.c_double * 3)(*kk), phase, fact, (ctypes.c_double * 3)(*coef1),
                   (ctypes.c_double * 3)(*coef2), (ctypes.c_double * 3)(*fda), k1_ff, *xyz0, target)


def make_params(n, lmagnetic=False, lentropy=False, nu=5e-3, eta=0.0, chi=0.0,
                cs0=1.0, gamma=1.0, cp=1.0, rho0=1.0, mu0=1.0, cdt=0.9, cdtv=0.25,
                dt=1e-3, L=2 * math.pi, lupw_uu=False, lupw_lnrho=False, gravz=0.0, bcz=None, lupw_ss=False,
                omega=0.0, theta=0.0, layers=None, z0=None):
    """Run parameters with the reference defaults (cdata.f90:235-239,
    eos_idealgas.f90:40-46).  bcz: None (periodic z) or one run.in entry per
    variable ('s', 'a', 'a2', 'c1', 'cT', or 'bottom:top').  layers: the
    conv-slab fields (layer_defaults).  z0: z of the first grid point
    (default -Lz/2)."""
    nx, ny, nz = (n, n, n) if np.isscalar(n) else n
    Lx, Ly, Lz = (L, L, L) if np.isscalar(L) else L
    nv = 4 + int(lentropy) + 3 * int(lmagnetic)
    bot, top = bc_codes(bcz, nv) if bcz is not None else ((ctypes.c_int * 8)(), (ctypes.c_int * 8)())
    p = Params(nx, ny, nz, int(lmagnetic), int(lentropy), Lx, Ly, Lz, nu, eta, chi,
               cs0, gamma, cp, rho0, mu0, cdt, cdtv, dt, int(lupw_uu), int(lupw_lnrho),
               gravz, int(bcz is None), bot, top, int(lupw_ss))
    p.omega, p.theta = omega, theta
    p.z0 = -Lz / 2 if z0 is None else z0
    for k, v in (layers or {}).items():
        if k not in LAYER_FIELDS:
            raise KeyError(k)
        setattr(p, k, v)
    return p


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        dp = ctypes.POINTER(ctypes.c_double)
        pp = ctypes.POINTER(Params)
        L.pco_nvar.argtypes = [pp]
        L.pco_iss.argtypes = [pp]
        L.pco_iax.argtypes = [pp]
        L.pco_eos_constants.argtypes = [pp, dp]
        L.pco_fill_ghosts.argtypes = [dp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        for name in ("pco_der", "pco_der2"):
            getattr(L, name).argtypes = [dp, pp, ctypes.c_int, ctypes.c_int, dp]
        L.pco_derij.argtypes = [dp, pp, ctypes.c_int, ctypes.c_int, ctypes.c_int, dp]
        L.pco_pde.argtypes = [dp, dp, pp, ctypes.c_int, dp, dp, ctypes.c_int]
        L.pco_advance.argtypes = [dp, dp, pp, ctypes.c_int, dp, dp, dp, ctypes.c_int]
        L.pco_addforce.argtypes = [dp, pp, ctypes.POINTER(Forcing)]
        L.pco_boundconds.argtypes = [dp, pp]
        L.pco_layer_profile.argtypes = [pp, ctypes.c_int, dp]
        L.pco_layer_profile.restype = None
        for name in ("pco_der", "pco_der2", "pco_derij", "pco_fill_ghosts", "pco_pde",
                     "pco_advance", "pco_eos_constants", "pco_addforce", "pco_boundconds"):
            getattr(L, name).restype = None
        _lib = L
    return _lib


def _ptr(a):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def nvar(p):
    return lib().pco_nvar(ctypes.byref(p))


def iss(p):
    return lib().pco_iss(ctypes.byref(p))


def iax(p):
    return lib().pco_iax(ctypes.byref(p))


def eos_constants(p):
    out = np.zeros(8)
    lib().pco_eos_constants(ctypes.byref(p), _ptr(out))
    keys = ("cs20", "gamma_m1", "gamma1", "cp1", "cv", "cv1", "lnrho0", "lnTT0")
    return dict(zip(keys, out.tolist()))


def with_ghosts(interior):
    """(nvar, nz, ny, nx) interior -> (nvar, mz, my, mx) with periodic ghosts."""
    interior = np.asarray(interior, dtype=np.float64)
    nv, nz, ny, nx = interior.shape
    g = NGHOST
    f = np.zeros((nv, nz + 2 * g, ny + 2 * g, nx + 2 * g))
    f[:, g:g + nz, g:g + ny, g:g + nx] = interior
    lib().pco_fill_ghosts(_ptr(f), nx, ny, nz, nv)
    return f


def interior(f):
    g = NGHOST
    return f[:, g:-g, g:-g, g:-g]


def der(f, p, k, j):
    out = np.zeros((p.nz, p.ny, p.nx))
    lib().pco_der(_ptr(f), ctypes.byref(p), k, j, _ptr(out))
    return out


def der2(f, p, k, j):
    out = np.zeros((p.nz, p.ny, p.nx))
    lib().pco_der2(_ptr(f), ctypes.byref(p), k, j, _ptr(out))
    return out


def derij(f, p, k, i, j):
    out = np.zeros((p.nz, p.ny, p.nx))
    lib().pco_derij(_ptr(f), ctypes.byref(p), k, i, j, _ptr(out))
    return out


def rhs(interior_f, p, want_dt=False, nthreads=0):
    """One RHS evaluation (df starting at zero).  Returns (rhs_interior,
    dt1_max, (urms, umax, brms))."""
    f = with_ghosts(interior_f)
    df = np.zeros_like(f)
    dt1 = np.zeros(1)
    diag = np.zeros(3)
    lib().pco_pde(_ptr(f), _ptr(df), ctypes.byref(p), int(want_dt), _ptr(dt1), _ptr(diag), nthreads)
    return interior(df).copy(), float(dt1[0]), tuple(diag.tolist())


def advance(interior_f, p, nsteps, t=0.0, dt=0.0, diag=False, nthreads=0, forcing=None):
    """nsteps full RK3 steps.  Returns (f_interior, t, dt, diag_series).
    forcing (a Forcing) is added after every step, as run.f90:575 calls
    addforce after time_step."""
    f = with_ghosts(interior_f)
    df = np.zeros_like(f)
    tt = ctypes.c_double(t)
    dd = ctypes.c_double(dt if p.dt == 0.0 else p.dt)
    series = np.zeros((max(nsteps, 1), 3)) if diag else None
    if forcing is None:
        lib().pco_advance(_ptr(f), _ptr(df), ctypes.byref(p), nsteps, ctypes.byref(tt), ctypes.byref(dd),
                          _ptr(series) if diag else None, nthreads)
    else:
        for it in range(nsteps):
            row = series[it:it + 1] if diag else None
            lib().pco_advance(_ptr(f), _ptr(df), ctypes.byref(p), 1, ctypes.byref(tt), ctypes.byref(dd),
                              _ptr(row) if diag else None, nthreads)
            lib().pco_addforce(_ptr(f), ctypes.byref(p), ctypes.byref(forcing))
    return interior(f).copy(), tt.value, dd.value, series


def addforce(interior_f, p, forcing):
    """The forcing of one step added to an interior state (copy)."""
    f = with_ghosts(interior_f)
    lib().pco_addforce(_ptr(f), ctypes.byref(p), ctypes.byref(forcing))
    return interior(f).copy()


def boundconds(f, p):
    """The run's boundary conditions applied in place to an f-array with
    ghost zones (nvar, mz, my, mx): periodic x and y, periodic or bcz in z."""
    lib().pco_boundconds(_ptr(f), ctypes.byref(p))
    return f


def layer_profile(p):
    """(nz, 4) array of {hcond, dlnhcond/dz, profz_heat, profz_cool} per z plane."""
    out = np.zeros((p.nz, 4))
    for n in range(p.nz):
        lib().pco_layer_profile(ctypes.byref(p), n, _ptr(out[n]))
    return out
