"""Parity of non-periodic z boundaries ('s', 'a', 'a2' of bc_sym_z,
boundcond.f90:3842-3900) and constant vertical gravity (gravity_simple,
gravz_profile='const') on the HIP path with the CPU oracle, through the
libpcmx C ABI: FP64 relative L-inf per variable <= 1e-12 after one step and
<= 1e-9 after 20, the mixed form <= 1e-5 after 100; kernels 1 and 3 (32x8
and 16x16 tiles), fixed and Courant steps.  The z ghost zones of the copied
f-array follow the rules applied to the copied state."""
import math

import numpy as np
import pytest

import oracle as O
import pcmx
import pcmx_ics
from test_gpu_parity import FP64_1STEP, FP64_100STEP, MIXED_100STEP, announce, rel_linf

pytestmark = pytest.mark.gpu

HYDRO_BC = ["s", "s", "a", "a2"]
MHD_BC = ["s", "s", "a", "a2:s", "a", "a", "s"]
ENT_BC = ["s", "s", "a", "a2", "s", "a", "a", "s"]


def state(n, lmag, lent, seed, gz, cs0=1.0):
    f0 = pcmx_ics.config_state("shells" if lmag else "noise", n, seed, lmagnetic=lmag, lentropy=lent)
    nx, ny, nz = (n, n, n) if np.isscalar(n) else n
    z = np.linspace(0.0, 1.0, nz)[:, None, None]
    f0[3] += 0.2 + gz / cs0 ** 2 * z          # a stratified density
    return f0


def both(n, bcz, lmag=False, lent=False, kernel=3, gz=-1.0, ldt=False, dt=1e-3, steps=1, seed=4,
         precision=pcmx.PREC_FP64, Lz=1.0):
    L = (2 * math.pi, 2 * math.pi * n[1] / n[0], Lz)
    gp = pcmx.make_params(n, lmagnetic=lmag, lentropy=lent, nu=4e-3, eta=3e-3, chi=2e-3,
                          gamma=5 / 3 if lent else 1.0, ldt=ldt, cdt=0.4, kernel=kernel, L=L, gravz=gz,
                          bcz=bcz, precision=precision)
    op = O.make_params(n, lmagnetic=lmag, lentropy=lent, nu=4e-3, eta=3e-3, chi=2e-3,
                       gamma=5 / 3 if lent else 1.0, cdt=0.4, dt=0.0 if ldt else dt, L=L, gravz=gz, bcz=bcz)
    f0 = state(n, lmag, lent, seed, gz)
    g, t, gdt = pcmx.advance(f0, gp, steps, dt)
    o, to, odt, _ = O.advance(f0, op, steps)
    return gp, op, f0, g, o, (t, gdt, to, odt)


@pytest.mark.parametrize("kernel", [1, 3])
@pytest.mark.parametrize("case", ["hydro", "mhd"])
def test_one_and_twenty_steps(case, kernel):
    n = (32, 16, 24)
    lmag = case == "mhd"
    bcz = MHD_BC if lmag else HYDRO_BC
    for steps, tol in ((1, FP64_1STEP), (20, FP64_100STEP)):
        announce(f"test_one_and_twenty_steps[{case}, kernel {kernel}] n={n} steps={steps}")
        gp, op, f0, g, o, _ = both(n, bcz, lmag=lmag, kernel=kernel, steps=steps)
        err = rel_linf(g, o)
        announce(f"  rel L-inf per variable: {err}")
        assert np.all(err <= tol), err
    # the boundaries matter: the periodic answer is far away
    per = O.advance(f0, O.make_params(n, lmagnetic=lmag, nu=4e-3, eta=3e-3, dt=1e-3,
                                      L=(2 * math.pi, math.pi, 24 / 23), gravz=-1.0), 20)[0]
    assert rel_linf(o, per).max() > 1e-6


def test_split_kernel_16x16_tiles():
    n = (16, 16, 20)
    announce(f"test_split_kernel_16x16_tiles: mhd n={n}")
    gp, op, f0, g, o, _ = both(n, MHD_BC, lmag=True, kernel=3, steps=3)
    err = rel_linf(g, o)
    announce(f"  rel L-inf per variable: {err}")
    assert np.all(err <= FP64_1STEP * 10), err


@pytest.mark.parametrize("kernel", [1, 3])
def test_entropy_courant_steps(kernel):
    n = (32, 16, 16)
    announce(f"test_entropy_courant_steps[kernel {kernel}] n={n}")
    gp, op, f0, g, o, (t, gdt, to, odt) = both(n, ENT_BC, lmag=True, lent=True, kernel=kernel, ldt=True,
                                                steps=3, dt=0.0)
    err = rel_linf(g, o)
    announce(f"  rel L-inf per variable: {err}, dt {gdt} vs {odt}")
    assert abs(gdt - odt) <= 1e-13 * odt
    assert np.all(err <= FP64_1STEP * 10), err


def test_mixed_form_hundred_steps():
    n = (32, 16, 24)
    announce(f"test_mixed_form_hundred_steps: hydro n={n}, kernel 3")
    gp, op, f0, g, o, _ = both(n, HYDRO_BC, kernel=3, steps=100, precision=pcmx.PREC_MIXED)
    err = rel_linf(g, o)
    announce(f"  mixed rel L-inf per variable: {err}")
    assert np.all(err <= MIXED_100STEP), err


def test_hydrostatic_atmosphere_stays_at_rest():
    n = (32, 8, 16)
    gz, cs0 = -1.3, 1.0
    announce("test_hydrostatic_atmosphere_stays_at_rest")
    gp = pcmx.make_params(n, nu=4e-3, kernel=3, L=(2.0, 0.5, 1.0), gravz=gz, bcz=HYDRO_BC)
    z = np.arange(n[2]) * gp.dz
    f0 = np.zeros((4, n[2], n[1], n[0]))
    f0[3] = 0.1 + gz / cs0 ** 2 * z[:, None, None]
    g, _, _ = pcmx.advance(f0, gp, 10, 1e-2)
    assert np.abs(g[:3]).max() < 1e-13
    assert np.abs(g[3] - f0[3]).max() < 1e-13


def test_copied_z_ghosts_follow_the_rules():
    n = (32, 16, 16)
    bcz = ["s", "a", "a2", "a2:s"]
    announce("test_copied_z_ghosts_follow_the_rules")
    gp = pcmx.make_params(n, kernel=3, L=(2 * math.pi, math.pi, 1.0), gravz=-0.5, bcz=bcz)
    f0 = state(n, False, False, 9, -0.5)
    f = pcmx.farray_from_interior(f0)
    pcmx.initialize_gpu(f, gp, nt=1)
    try:
        pcmx.time_step(2, 0.0, 1e-3)
        out = np.zeros_like(f)
        pcmx.copy_farray_from_gpu(out)
    finally:
        pcmx.finalize_gpu()
    cur = out[:4].copy()
    ref = O.with_ghosts(pcmx.interior_of(out, 4).copy())
    O.boundconds(ref, O.make_params(n, bcz=bcz))
    g = 3
    for plane in list(range(0, g)) + list(range(g + n[2], 2 * g + n[2])):
        np.testing.assert_array_equal(cur[:, plane], ref[:, plane])


def test_rejected_configurations():
    announce("test_rejected_configurations")
    with pytest.raises(pcmx.PcmxError, match="nz >= 8"):
        pcmx.initialize_gpu(None, pcmx.make_params((32, 8, 6), bcz=HYDRO_BC), nt=0)
    with pytest.raises(pcmx.PcmxError, match="kernel 2"):
        pcmx.initialize_gpu(None, pcmx.make_params((32, 8, 16), bcz=HYDRO_BC, kernel=2), nt=0)
    bad = pcmx.make_params((32, 8, 16), bcz=HYDRO_BC)
    bad.bcz_top[2] = 7
    with pytest.raises(pcmx.PcmxError, match="code 7"):
        pcmx.initialize_gpu(None, bad, nt=0)


def test_graph_replay_with_z_boundaries():
    from test_gpu_parity import _steps
    n = (32, 16, 24)
    announce("test_graph_replay_with_z_boundaries: mhd, kernel 3, 6 steps")
    gp = pcmx.make_params(n, lmagnetic=True, nu=4e-3, eta=3e-3, kernel=3, L=(2 * math.pi, math.pi, 1.0),
                          gravz=-1.0, bcz=MHD_BC)
    f0 = state(n, True, False, 4, -1.0)
    a = _steps(f0, gp, 6, 1e-3, timing=False, calls=2)
    b = _steps(f0, gp, 6, 1e-3, timing=True, calls=2)
    np.testing.assert_array_equal(a[0], b[0])


@pytest.mark.parametrize("kernel", [1, 3])
def test_forcing_with_z_boundaries(kernel):
    """Helical forcing on a grid whose z points include the boundaries
    (z = z0 + n dz): the phasors follow that grid."""
    import pcmx_forcing as PF
    n = (32, 16, 24)
    announce(f"test_forcing_with_z_boundaries[kernel {kernel}]")
    L = (2 * math.pi, math.pi, 1.0)
    gp = pcmx.make_params(n, lmagnetic=True, nu=4e-3, eta=3e-3, kernel=kernel, L=L, gravz=-1.0, bcz=MHD_BC)
    op = O.make_params(n, lmagnetic=True, nu=4e-3, eta=3e-3, dt=1e-3, L=L, gravz=-1.0, bcz=MHD_BC)
    pars = PF.helical_pars((1, 2, 3), 0.4, 1.3, 0.2, 1e-3, 3.0)
    xyz0 = (-math.pi, -0.5 * math.pi, -0.5)
    kw = dict(kk=pars["kk"], phase=pars["phase"], fact=pars["fact"], coef1=pars["coef1"], coef2=pars["coef2"],
              fda=pars["fda"], k1_ff=1.0, xyz0=xyz0, target=0)
    ofo = O.make_forcing(pars["kk"], pars["phase"], pars["fact"], pars["coef1"], pars["coef2"], pars["fda"],
                         k1_ff=1.0, xyz0=xyz0, target=0)
    f0 = state(n, True, False, 4, -1.0)
    g, _, _ = pcmx.advance(f0, gp, 5, 1e-3, forcing=kw)
    o, _, _, _ = O.advance(f0, op, 5, forcing=ofo)
    err = rel_linf(g, o)
    announce(f"  rel L-inf per variable: {err}")
    assert np.all(err <= FP64_1STEP * 10), err
    plain, _, _, _ = O.advance(f0, op, 5)
    assert rel_linf(o, plain).max() > 1e-6


def test_snapshot_and_dt1_with_z_boundaries(tmp_path):
    import pcmx_io as I
    n = (32, 16, 16)
    announce("test_snapshot_and_dt1_with_z_boundaries")
    L = (2 * math.pi, math.pi, 1.5)
    gp = pcmx.make_params(n, lmagnetic=True, nu=4e-3, eta=3e-3, kernel=3, L=L, gravz=-1.0, bcz=MHD_BC, ldt=True,
                          cdt=0.4)
    f0 = state(n, True, False, 6, -1.0)
    f = pcmx.farray_from_interior(f0)
    pcmx.initialize_gpu(f, gp, nt=1)
    try:
        dt1 = pcmx.gpu_prepare_for_first_substep(0.0)
        pcmx.time_step(2, 0.0, 0.0)
        path = tmp_path / "var.dat"
        xyz0 = (-math.pi, -0.5 * math.pi, -0.75)
        pcmx.write_snapshot(str(path), 0.0, xyz0)
        copy = np.zeros_like(f)
        pcmx.copy_farray_from_gpu(copy)
    finally:
        pcmx.finalize_gpu()
    op = O.make_params(n, lmagnetic=True, nu=4e-3, eta=3e-3, dt=0.0, L=L, gravz=-1.0, bcz=MHD_BC, cdt=0.4)
    _, dt1o, _ = O.rhs(f0, op, want_dt=True)
    assert abs(dt1 - dt1o) <= 1e-13 * dt1o
    gvar, t, x, y, z, dx, dy, dz = I.read_var(str(path), 7, n)
    np.testing.assert_array_equal(gvar, copy[:7])
    # boundary points on the first and last interior z plane
    assert z[3] == xyz0[2] and abs(z[3 + n[2] - 1] - (xyz0[2] + L[2])) < 1e-14
    assert dz == gp.dz == L[2] / (n[2] - 1)


def test_per_substep_api_with_z_boundaries_and_courant_dt():
    """The reference's call sequence (rhs_gpu per substep) equals whole steps,
    bitwise, with z boundaries, gravity and a Courant dt."""
    n = (32, 16, 16)
    announce("test_per_substep_api_with_z_boundaries_and_courant_dt")
    gp = pcmx.make_params(n, lmagnetic=True, nu=4e-3, eta=3e-3, kernel=3, L=(2 * math.pi, math.pi, 1.0),
                          gravz=-1.0, bcz=MHD_BC, ldt=True, cdt=0.4)
    f0 = state(n, True, False, 8, -1.0)
    a, ta, dta = pcmx.advance(f0, gp, 3, 0.0, per_substep=True)
    b, tb, dtb = pcmx.advance(f0, gp, 3, 0.0)
    np.testing.assert_array_equal(a, b)
    assert ta == tb and dta == dtb
