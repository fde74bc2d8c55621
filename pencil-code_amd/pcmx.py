"""Python host mirror of the Pencil Code ``module GPU`` interface over libpcmx.

The reference's Fortran module (src/gpu.h, src/gpu_astaroth.f90) exposes
register_GPU, initialize_GPU, load_farray_to_GPU, before_boundary_gpu,
rhs_GPU, update_after_substep_gpu, copy_farray_from_GPU,
gpu_prepare_for_first_substep, get_gpu_reduced_vars, update_on_gpu,
test_gpu_bcs, split_update_gpu and finalize_GPU.  This module exposes the
same names with the same argument meaning over the C ABI of
``include/pcmx.h`` (loaded with ctypes), so Python drivers and the tests
read like the reference's call sequence in equ.f90 / timestep.f90.

The f-array is the reference layout ``f(mx,my,mz,mfarray)`` in Fortran
order, i.e. a C-ordered numpy array of shape ``(mfarray, mz, my, mx)``.

There is no CPU fallback: if libpcmx.so is missing or no GPU is present,
every call raises.
"""
from __future__ import annotations

import ctypes
import math
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libpcmx.so")
NGHOST = 3
N_GPU_REDUCED_VARS = 20

PREC_FP64 = 0
PREC_MIXED = 1
KERNEL_AUTO, KERNEL_DIRECT, KERNEL_ZMARCH, KERNEL_ZSPLIT = 0, 1, 2, 3


class PcmxError(RuntimeError):
    pass


class Params(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
        ("mvar", ctypes.c_int),
        ("iux", ctypes.c_int), ("ilnrho", ctypes.c_int), ("iss", ctypes.c_int), ("iax", ctypes.c_int),
        ("dx", ctypes.c_double), ("dy", ctypes.c_double), ("dz", ctypes.c_double),
        ("nu", ctypes.c_double), ("eta", ctypes.c_double), ("chi", ctypes.c_double),
        ("cs0", ctypes.c_double), ("gamma", ctypes.c_double), ("cp", ctypes.c_double),
        ("rho0", ctypes.c_double), ("mu0", ctypes.c_double),
        ("cdt", ctypes.c_double), ("cdtv", ctypes.c_double),
        ("ldt", ctypes.c_int), ("precision", ctypes.c_int), ("kernel", ctypes.c_int),
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
        ("z0", ctypes.c_double), ("lxyz", ctypes.c_double * 3),
    ]


BC_CODES = {"s": 1, "a": 2, "a2": 3, "c1": 4, "cT": 5}  # PCMX_BC_S, _A, _A2, _C1, _CT

# layer fields of Params (samples/conv-slab physics), set through make_params(layers=...)
LAYER_FIELDS = ("lheatc_kprof", "cooling_profile", "hcond0", "hcond1", "hcond2", "widthss", "z1", "z2",
                "luminosity", "wheat", "cool", "wcool", "cs2cool", "fbot", "ftop", "cs2bot", "cs2top")


def layer_defaults(gamma, cp, gravz, bcz_ss, hcond0=0.0, mpoly0=1.5, mpoly1=1.5, mpoly2=1.5, widthss=0.05,
                   z1=0.0, z2=1.0, luminosity=0.0, wheat=0.1, cool=0.0, wcool=0.1, cs2cool=0.0,
                   cooling_profile="gaussian", cs2top=None, cs2bot=None, fbot=None, ftop=None, kprof=True):
    """Layer fields derived as initialize_energy does (entropy.f90:805-890):
    hcond1 = (mpoly1+1)/(mpoly0+1), hcond2 = (mpoly2+1)/(mpoly0+1) (the
    K-profile of a piecewise polytrope); Fbot (Ftop) =
    -gamma/(cp*(gamma-1))*hcond0*gravz/(mpoly0+1) when the bottom (top) rule
    of ss is 'c1', else 0, unless given; with cooling, cs2top =
    (cool*cs2cool + cool2*cs2cool2)/(cool+cool2) with cool2 = cs2cool2 = 0,
    unless given.  bcz_ss: the run.in entry of ss, e.g. 'c1:cT'."""
    lo, _, hi = bcz_ss.partition(":")
    hi = hi or lo
    flux = -gamma / (cp * (gamma - 1.0)) * hcond0 * gravz / (mpoly0 + 1)
    if cs2top is None:
        cs2top = (cool * cs2cool + 0.0 * 0.0) / (cool + 0.0) if cool != 0.0 else 0.0
    return dict(lheatc_kprof=int(kprof), cooling_profile={"gaussian": 0, "step": 1}[cooling_profile],
                hcond0=hcond0, hcond1=(mpoly1 + 1.) / (mpoly0 + 1.), hcond2=(mpoly2 + 1.) / (mpoly0 + 1.),
                widthss=widthss, z1=z1, z2=z2, luminosity=luminosity, wheat=wheat, cool=cool, wcool=wcool,
                cs2cool=cs2cool, fbot=(flux if lo == "c1" else 0.0) if fbot is None else fbot,
                ftop=(flux if hi == "c1" else 0.0) if ftop is None else ftop,
                cs2bot=0.0 if cs2bot is None else cs2bot, cs2top=cs2top)


def bc_codes(bcz, nv):
    """Per-variable (bottom, top) code arrays from run.in style entries:
    's', 'a', 'a2', or 'bottom:top' such as 'a2:s'."""
    if len(bcz) != nv:
        raise ValueError(f"{len(bcz)} z boundary conditions for {nv} variables")
    bot, top = (ctypes.c_int * 8)(), (ctypes.c_int * 8)()
    for k, b in enumerate(bcz):
        lo, _, hi = b.partition(":")
        bot[k], top[k] = BC_CODES[lo], BC_CODES[hi or lo]
    return bot, top


class Box(ctypes.Structure):
    """One box of an ensemble (pcmx_box)."""
    _fields_ = [
        ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
        ("dx", ctypes.c_double), ("dy", ctypes.c_double), ("dz", ctypes.c_double),
        ("nu", ctypes.c_double), ("eta", ctypes.c_double), ("dt", ctypes.c_double),
    ]


class Forcing(ctypes.Structure):
    """Helical forcing parameters (pcmx_forcing)."""
    _fields_ = [
        ("active", ctypes.c_int), ("target", ctypes.c_int), ("scale_sqrt_dt", ctypes.c_int),
        ("reserved", ctypes.c_int),
        ("kk", ctypes.c_double * 3), ("phase", ctypes.c_double), ("fact", ctypes.c_double),
        ("coef1", ctypes.c_double * 3), ("coef2", ctypes.c_double * 3), ("fda", ctypes.c_double * 3),
        ("k1_ff", ctypes.c_double), ("xyz0", ctypes.c_double * 3),
    ]


ENS_NO_GRAPH = 1


def make_params(n, lmagnetic=False, lentropy=False, nu=5e-3, eta=0.0, chi=0.0, cs0=1.0, gamma=1.0,
                cp=1.0, rho0=1.0, mu0=1.0, cdt=0.9, cdtv=0.25, ldt=False, precision=PREC_FP64,
                kernel=None, L=2 * math.pi, lupw_uu=False, lupw_lnrho=False, gravz=0.0, bcz=None,
                lupw_ss=False, omega=0.0, theta=0.0, layers=None, z0=None):
    """Parameters of one periodic box with the reference defaults
    (cdata.f90:235-239, eos_idealgas.f90:40-46).  Slots follow the
    registration order uu, lnrho, [ss], [aa] (register.f90:137-146).
    kernel None takes PCMX_KERNEL from the environment (0 auto, 1 direct,
    2 z-marching, 3 z-marching with the flow and field halves on separate
    wavefronts), default auto.  bcz: None for a periodic z, else one run.in
    entry per variable ('s', 'a', 'a2', 'c1', 'cT', 'bottom:top'); the z
    spacing is then Lz/(nz-1), boundary points on the first and last plane.
    omega, theta: rotation; layers: the conv-slab fields (layer_defaults);
    z0: z of the first grid point (default -Lz/2)."""
    if kernel is None:
        kernel = int(os.environ.get("PCMX_KERNEL", KERNEL_AUTO))
    nx, ny, nz = (n, n, n) if np.isscalar(n) else n
    Lx, Ly, Lz = (L, L, L) if np.isscalar(L) else L
    iss = 4 if lentropy else -1
    iax = (5 if lentropy else 4) if lmagnetic else -1
    mvar = 4 + (1 if lentropy else 0) + (3 if lmagnetic else 0)
    bot, top = bc_codes(bcz, mvar) if bcz is not None else ((ctypes.c_int * 8)(), (ctypes.c_int * 8)())
    dz = Lz / nz if bcz is None else Lz / (nz - 1)
    p = Params(nx, ny, nz, mvar, 0, 3, iss, iax, Lx / nx, Ly / ny, dz, nu, eta, chi, cs0, gamma,
               cp, rho0, mu0, cdt, cdtv, int(ldt), precision, kernel, int(lupw_uu), int(lupw_lnrho),
               gravz, int(bcz is None), bot, top, int(lupw_ss))
    p.omega, p.theta = omega, theta
    p.z0 = -Lz / 2 if z0 is None else z0
    p.lxyz[:] = (Lx, Ly, Lz)
    for k, v in (layers or {}).items():
        if k not in LAYER_FIELDS:
            raise KeyError(k)
        setattr(p, k, v)
    return p


_lib = None


def lib():
    """Load libpcmx.so; raise if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise PcmxError(f"{LIB_PATH} is missing: build it with __graft_entry__.build()")
        L = ctypes.CDLL(LIB_PATH)
        dp = ctypes.POINTER(ctypes.c_double)
        L.pcmx_last_error.restype = ctypes.c_char_p
        L.pcmx_last_error.argtypes = []
        sig = {
            "pcmx_register": [],
            "pcmx_initialize": [dp, ctypes.c_int, ctypes.POINTER(Params), ctypes.c_double, ctypes.c_int],
            "pcmx_finalize": [],
            "pcmx_load_farray": [dp, ctypes.c_int],
            "pcmx_copy_farray": [dp, ctypes.c_int],
            "pcmx_before_boundary": [ctypes.c_int, ctypes.c_int, ctypes.c_double],
            "pcmx_rhs": [ctypes.c_int, ctypes.c_double, dp],
            "pcmx_update_after_substep": [],
            "pcmx_prepare_for_first_substep": [ctypes.c_double, dp],
            "pcmx_get_reduced_vars": [dp],
            "pcmx_update_by_name": [ctypes.c_char_p, ctypes.c_double],
            "pcmx_update_by_index": [ctypes.c_int, ctypes.c_double],
            "pcmx_test_bcs": [dp],
            "pcmx_split_update": [],
            "pcmx_time_step": [ctypes.c_int, dp, dp, ctypes.c_int],
            "pcmx_kernel_timing": [dp],
            "pcmx_tune": [dp],
            "pcmx_set_timing": [ctypes.c_int],
            "pcmx_ens_create": [ctypes.POINTER(Params), ctypes.POINTER(Box), ctypes.c_int, ctypes.c_int],
            "pcmx_ens_destroy": [],
            "pcmx_ens_load": [ctypes.c_int, dp],
            "pcmx_ens_copy": [ctypes.c_int, dp],
            "pcmx_ens_step": [ctypes.c_int, ctypes.c_int],
            "pcmx_ens_reduced": [dp],
            "pcmx_ens_tune": [ctypes.c_int],
            "pcmx_ens_set_timing": [ctypes.c_int],
            "pcmx_ens_timing": [dp],
            "pcmx_write_snapshot": [ctypes.c_char_p, ctypes.c_double, dp],
            "pcmx_set_forcing": [ctypes.POINTER(Forcing)],
            "pcmx_write_record": [ctypes.c_char_p, ctypes.c_int, ctypes.POINTER(dp), ctypes.POINTER(ctypes.c_long),
                                  ctypes.c_int, ctypes.c_long],
        }
        for name, args in sig.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = ctypes.c_int
        _lib = L
    return _lib


def _check(rc):
    if rc < 0:
        raise PcmxError(lib().pcmx_last_error().decode())
    return rc


def _ptr(a):
    if a.dtype != np.float64 or not a.flags["C_CONTIGUOUS"]:
        raise PcmxError("f-array must be a C-contiguous float64 array of shape (mfarray, mz, my, mx)")
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


# ------------------------------------------------------------------ module GPU
def register_gpu():
    _check(lib().pcmx_register())


def initialize_gpu(f, par, t=0.0, nt=1):
    """initialize_GPU(f): allocate device state; load f when nt > 0."""
    if f is not None:
        _check(lib().pcmx_initialize(_ptr(f), f.shape[0], ctypes.byref(par), t, nt))
    else:
        _check(lib().pcmx_initialize(None, 0, ctypes.byref(par), t, 0))


def finalize_gpu():
    _check(lib().pcmx_finalize())


def load_farray_to_gpu(f):
    _check(lib().pcmx_load_farray(_ptr(f), f.shape[0]))


def copy_farray_from_gpu(f):
    _check(lib().pcmx_copy_farray(_ptr(f), f.shape[0]))


def before_boundary_gpu(lrmv, isubstep, t):
    _check(lib().pcmx_before_boundary(int(lrmv), isubstep, t))


def rhs_gpu(isubstep, t, dt):
    """rhs_GPU: one RK substep.  Returns dt (the Courant step on substep 1
    of a Courant-limited run, else the dt passed in)."""
    d = ctypes.c_double(dt)
    _check(lib().pcmx_rhs(isubstep, t, ctypes.byref(d)))
    return d.value


def update_after_substep_gpu():
    _check(lib().pcmx_update_after_substep())


def split_update_gpu():
    _check(lib().pcmx_split_update())


def gpu_prepare_for_first_substep(t=0.0):
    d = ctypes.c_double(0.0)
    _check(lib().pcmx_prepare_for_first_substep(t, ctypes.byref(d)))
    return d.value


def get_gpu_reduced_vars():
    dst = np.zeros(N_GPU_REDUCED_VARS)
    _check(lib().pcmx_get_reduced_vars(_ptr(dst)))
    return dst


def update_on_gpu(index, varname=None, value=None):
    """update_on_gpu(index, varname, value): index < 0 looks the name up and
    returns its index; an unknown name raises (the reference: fatal_error)."""
    if index >= 0:
        _check(lib().pcmx_update_by_index(index, value))
        return index
    return _check(lib().pcmx_update_by_name(varname.encode(), value))


def test_gpu_bcs():
    d = ctypes.c_double(-1.0)
    _check(lib().pcmx_test_bcs(ctypes.byref(d)))
    return d.value


def time_step(nsteps, t, dt, ldiag=True):
    """nsteps full RK3 steps on the device.  Returns (t, dt)."""
    tt, dd = ctypes.c_double(t), ctypes.c_double(dt)
    _check(lib().pcmx_time_step(nsteps, ctypes.byref(tt), ctypes.byref(dd), int(ldiag)))
    return tt.value, dd.value


def set_forcing(kk, phase, fact, coef1, coef2, fda=(1.0, 1.0, 1.0), k1_ff=1.0,
                xyz0=(-math.pi, -math.pi, -math.pi), target=0, scale_sqrt_dt=False):
    """Helical forcing added after every following step (pcmx_set_forcing):
    the parameters forcing_pars_hel returns (forcing.f90:1731-1746), k1_ff,
    the box origin and the 0-based slot of the forced vector (iux, or iax
    for lmagnetic_forcing).  With scale_sqrt_dt, fact is force/ffnorm and
    each step multiplies it by sqrt(dt)."""
    v3 = lambda a: (ctypes.c_double * 3)(*[float(x) for x in a])  # noqa: E731
    fr = Forcing(1, int(target), int(bool(scale_sqrt_dt)), 0, v3(kk), float(phase), float(fact), v3(coef1),
                 v3(coef2), v3(fda), float(k1_ff), v3(xyz0))
    _check(lib().pcmx_set_forcing(ctypes.byref(fr)))


def clear_forcing():
    _check(lib().pcmx_set_forcing(ctypes.byref(Forcing())))


def write_snapshot(path, t, xyz0=(-math.pi, -math.pi, -math.pi)):
    """Writes the device state as a var.dat snapshot (the reference's
    output_snap layout for one process, src/io_dist.f90:108-199) without
    copying the f-array to the host.  xyz0: the box origin (start.in xyz0,
    default -pi in every direction)."""
    o = (ctypes.c_double * 3)(*[float(v) for v in xyz0])
    _check(lib().pcmx_write_snapshot(os.fsencode(path), float(t), o))


def write_record(path, parts, append=False, max_subrecord=0):
    """One Fortran unformatted sequential record (gfortran's convention)
    holding the given float64 arrays back to back; max_subrecord (bytes, 0 =
    the runtime's limit) exists for tests of the subrecord layout."""
    arrs = [np.ascontiguousarray(a, dtype=np.float64).ravel() for a in parts]
    dp = ctypes.POINTER(ctypes.c_double)
    ptrs = (dp * max(1, len(arrs)))(*[a.ctypes.data_as(dp) for a in arrs])
    counts = (ctypes.c_long * max(1, len(arrs)))(*[a.size for a in arrs])
    _check(lib().pcmx_write_record(os.fsencode(path), int(bool(append)), ptrs, counts, len(arrs), int(max_subrecord)))


def tune():
    """pcmx_tune: pick the fastest tile shape and z-chunk length for the
    loaded state.  Returns (ms per launch, tile width, z-chunk length)."""
    out = np.zeros(3)
    _check(lib().pcmx_tune(_ptr(out)))
    return float(out[0]), int(out[1]), int(out[2])


def set_timing(enable):
    _check(lib().pcmx_set_timing(int(enable)))


def kernel_timing():
    out = np.zeros(2)
    _check(lib().pcmx_kernel_timing(_ptr(out)))
    return float(out[0]), int(out[1])


# ------------------------------------------------------------------ helpers
def farray_from_interior(interior, mfarray=None):
    """(mvar, nz, ny, nx) -> zero-ghost f-array (mfarray, mz, my, mx)."""
    interior = np.asarray(interior, dtype=np.float64)
    mv, nz, ny, nx = interior.shape
    mf = mfarray or mv
    f = np.zeros((mf, nz + 2 * NGHOST, ny + 2 * NGHOST, nx + 2 * NGHOST))
    g = NGHOST
    f[:mv, g:g + nz, g:g + ny, g:g + nx] = interior
    return f


def interior_of(f, mvar):
    g = NGHOST
    return f[:mvar, g:-g, g:-g, g:-g]


def advance(interior, par, nsteps, dt, t=0.0, per_substep=False, forcing=None):
    """Convenience driver in the reference's call order: load, nsteps steps,
    copy back.  Returns (interior_after, t, dt).  forcing: keyword
    arguments of set_forcing, applied after every step."""
    f = farray_from_interior(interior)
    initialize_gpu(f, par, t=t, nt=1)
    try:
        if forcing:
            set_forcing(**forcing)
        if per_substep:
            for _ in range(nsteps):
                ds = 0.0
                for isub in (1, 2, 3):
                    before_boundary_gpu(False, isub, t)
                    dt = rhs_gpu(isub, t, dt)
                    ds = 1.0 if isub == 1 else (-5.0 / 9.0 if isub == 2 else -153.0 / 128.0) * ds + 1.0
                    t = t + ds * (dt * (1.0 / 3.0, 15.0 / 16.0, 8.0 / 15.0)[isub - 1])
                    update_after_substep_gpu()
        else:
            t, dt = time_step(nsteps, t, dt)
        out = np.zeros_like(f)
        copy_farray_from_gpu(out)
        return interior_of(out, par.mvar).copy(), t, dt
    finally:
        finalize_gpu()


# ------------------------------------------------------------------ ensembles
def make_box(n, nu, eta, dt, L=2 * math.pi):
    nx, ny, nz = (n, n, n) if np.isscalar(n) else n
    Lx, Ly, Lz = (L, L, L) if np.isscalar(L) else L
    return Box(nx, ny, nz, Lx / nx, Ly / ny, Lz / nz, nu, eta, dt)


def ens_create(par, boxes, flags=0):
    """pcmx_ens_create: boxes is a list of Box; par gives the shared physics."""
    arr = (Box * len(boxes))(*boxes)
    _check(lib().pcmx_ens_create(ctypes.byref(par), arr, len(boxes), flags))


def ens_destroy():
    _check(lib().pcmx_ens_destroy())


def ens_load(ibox, interior):
    a = np.ascontiguousarray(interior, dtype=np.float64)
    _check(lib().pcmx_ens_load(ibox, _ptr(a)))


def ens_copy(ibox, shape):
    out = np.zeros(shape)
    _check(lib().pcmx_ens_copy(ibox, _ptr(out)))
    return out


def ens_step(nsteps, ldiag=True):
    _check(lib().pcmx_ens_step(nsteps, int(ldiag)))


def ens_reduced(nbox):
    out = np.zeros(3 * nbox)
    _check(lib().pcmx_ens_reduced(_ptr(out)))
    return out.reshape(nbox, 3)


def ens_tune(nsteps_probe=2):
    _check(lib().pcmx_ens_tune(nsteps_probe))


def ens_set_timing(enable):
    _check(lib().pcmx_ens_set_timing(int(enable)))


def ens_timing():
    """(mean ms per substep, substeps, tile width, z-chunk limit, blocks per CU, work items)."""
    out = np.zeros(6)
    _check(lib().pcmx_ens_timing(_ptr(out)))
    return float(out[0]), int(out[1]), int(out[2]), int(out[3]), int(out[4]), int(out[5])
