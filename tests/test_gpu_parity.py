"""Parity of the HIP path (through the libpcmx C ABI) with the CPU oracle.

Tolerances are the north-star ones: FP64 relative L-inf per variable
<= 1e-12 after one RK3 step and <= 1e-9 after 100 steps; the mixed form
<= 1e-5 per variable after 100 steps.  Relative L-inf of a variable is
max|gpu - oracle| / max|oracle|."""
import math
import os
import sys

import numpy as np
import pytest

import oracle as O
import pcmx
import pcmx_ics

pytestmark = pytest.mark.gpu

FP64_1STEP = 1e-12
FP64_100STEP = 1e-9
MIXED_100STEP = 1e-5


def rel_linf(a, b):
    out = []
    for v in range(a.shape[0]):
        den = np.abs(b[v]).max()
        num = np.abs(a[v] - b[v]).max()
        out.append(num / den if den > 0 else num)
    return np.array(out)


def oracle_params(gp, dt):
    ent, mag = gp.iss >= 0, gp.iax >= 0
    return O.make_params((gp.nx, gp.ny, gp.nz), lmagnetic=mag, lentropy=ent, nu=gp.nu, eta=gp.eta,
                         chi=gp.chi, cs0=gp.cs0, gamma=gp.gamma, cp=gp.cp, rho0=gp.rho0, mu0=gp.mu0,
                         cdt=gp.cdt, cdtv=gp.cdtv, dt=0.0 if gp.ldt else dt,
                         L=(gp.dx * gp.nx, gp.dy * gp.ny, gp.dz * gp.nz),
                         lupw_uu=bool(gp.lupw_uu), lupw_lnrho=bool(gp.lupw_lnrho), lupw_ss=bool(gp.lupw_ss))


def announce(msg):
    print(msg, flush=True)
    sys.stdout.flush()


CASES = {
    # hydro, 4 vars, the CPU-runnable plain case
    "hydro32": dict(n=32, kind="noise", mag=False, ent=False, nu=5e-3, eta=0.0, chi=0.0, gamma=1.0,
                    dt=1e-3, ldt=False),
    # isothermal MHD, 7 vars, band-limited fields
    "mhd32": dict(n=32, kind="shells", mag=True, ent=False, nu=5e-3, eta=5e-3, chi=0.0, gamma=1.0,
                  dt=2e-3, ldt=False),
    # MHD + entropy, 8 vars, Courant-limited dt
    "mhdss24": dict(n=24, kind="noise", mag=True, ent=True, nu=2e-3, eta=2e-3, chi=2e-3, gamma=5 / 3,
                    dt=0.0, ldt=True, cdt=0.4),
    # non-cubic, unequal spacing
    "mhd_ragged": dict(n=(16, 24, 20), kind="shells", mag=True, ent=False, nu=5e-3, eta=4e-3, chi=0.0,
                       gamma=1.0, dt=1e-3, ldt=False, L=(2 * math.pi, 3.0, 2.5)),
    # odd sizes that do not fill a tile
    "hydro_odd": dict(n=(13, 11, 9), kind="noise", mag=False, ent=False, nu=1e-2, eta=0.0, chi=0.0,
                      gamma=1.0, dt=1e-3, ldt=False),
}


def setup_case(name, precision=pcmx.PREC_FP64, kernel=None, seed=7, lupw_uu=False, lupw_lnrho=False):
    c = CASES[name]
    gp = pcmx.make_params(c["n"], lmagnetic=c["mag"], lentropy=c["ent"], nu=c["nu"], eta=c["eta"],
                          chi=c["chi"], gamma=c["gamma"], ldt=c["ldt"], cdt=c.get("cdt", 0.9),
                          precision=precision, kernel=kernel, L=c.get("L", 2 * math.pi),
                          lupw_uu=lupw_uu, lupw_lnrho=lupw_lnrho)
    f0 = pcmx_ics.config_state(c["kind"], c["n"], seed, lmagnetic=c["mag"], lentropy=c["ent"])
    return c, gp, f0


def test_periodic_ghost_fill_matches_host():
    announce("test_periodic_ghost_fill_matches_host: 16x12x10 hydro")
    gp = pcmx.make_params((16, 12, 10))
    pcmx.initialize_gpu(None, gp, nt=0)
    try:
        assert pcmx.test_gpu_bcs() == 0.0
    finally:
        pcmx.finalize_gpu()


def test_load_copy_round_trip_fills_ghosts():
    announce("test_load_copy_round_trip_fills_ghosts: 12^3 mhd")
    c, gp, f0 = setup_case("mhd32")
    gp = pcmx.make_params(12, lmagnetic=True, nu=5e-3, eta=5e-3)
    f0 = pcmx_ics.config_state("noise", 12, 3, lmagnetic=True)
    f = pcmx.farray_from_interior(f0, mfarray=9)   # two extra (auxiliary) slots
    pcmx.initialize_gpu(f, gp, nt=1)
    try:
        out = np.full_like(f, 123.0)
        pcmx.copy_farray_from_gpu(out)
        np.testing.assert_array_equal(out[:7], O.with_ghosts(f0))
        assert np.all(out[7:] == 123.0)               # auxiliaries untouched
    finally:
        pcmx.finalize_gpu()


@pytest.mark.parametrize("name", list(CASES))
def test_one_step_fp64(name):
    c, gp, f0 = setup_case(name)
    announce(f"test_one_step_fp64[{name}]: n={c['n']} mag={c['mag']} ent={c['ent']} ldt={c['ldt']}")
    g1, t, dt = pcmx.advance(f0, gp, 1, c["dt"])
    o1, to, dto, _ = O.advance(f0, oracle_params(gp, c["dt"]), 1)
    err = rel_linf(g1, o1)
    announce(f"  rel L-inf per variable: {err}")
    assert np.all(err <= FP64_1STEP), err
    assert abs(t - to) <= 1e-15 * max(1.0, abs(to))
    if c["ldt"]:
        assert abs(dt - dto) <= 1e-13 * dto


@pytest.mark.parametrize("name", ["hydro32", "mhd32", "mhdss24"])
def test_hundred_steps_fp64(name):
    c, gp, f0 = setup_case(name)
    n = 100 if name != "mhdss24" else 60
    announce(f"test_hundred_steps_fp64[{name}]: {n} steps")
    g1, t, dt = pcmx.advance(f0, gp, n, c["dt"])
    o1, to, dto, _ = O.advance(f0, oracle_params(gp, c["dt"]), n)
    err = rel_linf(g1, o1)
    announce(f"  rel L-inf per variable: {err}")
    assert np.all(err <= FP64_100STEP), err


def test_per_substep_api_matches_time_step():
    c, gp, f0 = setup_case("mhdss24")
    announce("test_per_substep_api_matches_time_step: mhdss24, 3 steps")
    a, ta, dta = pcmx.advance(f0, gp, 3, c["dt"], per_substep=True)
    b, tb, dtb = pcmx.advance(f0, gp, 3, c["dt"], per_substep=False)
    np.testing.assert_array_equal(a, b)
    assert ta == tb and dta == dtb


def test_diagnostics_match_oracle():
    c, gp, f0 = setup_case("mhd32")
    announce("test_diagnostics_match_oracle: mhd32")
    f = pcmx.farray_from_interior(f0)
    pcmx.initialize_gpu(f, gp, nt=1)
    try:
        pcmx.rhs_gpu(1, 0.0, c["dt"])
        red = pcmx.get_gpu_reduced_vars()
    finally:
        pcmx.finalize_gpu()
    _, _, (urms, umax, brms) = O.rhs(f0, oracle_params(gp, c["dt"]))
    assert abs(red[0] - urms) <= 1e-14 * urms
    assert abs(red[1] - umax) <= 1e-14 * umax
    assert abs(red[2] - brms) <= 1e-14 * brms
    np.testing.assert_allclose(red[5:12], f0.reshape(7, -1).mean(axis=1), rtol=0, atol=1e-15)


def test_prepare_for_first_substep_dt1():
    c, gp, f0 = setup_case("mhdss24")
    announce("test_prepare_for_first_substep_dt1: mhdss24")
    f = pcmx.farray_from_interior(f0)
    pcmx.initialize_gpu(f, gp, nt=1)
    try:
        dt1 = pcmx.gpu_prepare_for_first_substep(0.0)
    finally:
        pcmx.finalize_gpu()
    _, dt1o, _ = O.rhs(f0, oracle_params(gp, 0.0), want_dt=True)
    assert abs(dt1 - dt1o) <= 1e-13 * dt1o


def test_mixed_form_error_after_100_steps():
    c, gp, f0 = setup_case("hydro32", precision=pcmx.PREC_MIXED)
    announce("test_mixed_form_error_after_100_steps: hydro32")
    g1, _, _ = pcmx.advance(f0, gp, 100, c["dt"])
    o1, _, _, _ = O.advance(f0, oracle_params(gp, c["dt"]), 100)
    err = rel_linf(g1, o1)
    announce(f"  mixed rel L-inf per variable: {err}")
    assert np.all(err <= MIXED_100STEP), err


def test_update_on_gpu_by_name_changes_viscosity():
    c, gp, f0 = setup_case("hydro32")
    announce("test_update_on_gpu_by_name_changes_viscosity")
    f = pcmx.farray_from_interior(f0)
    pcmx.initialize_gpu(f, gp, nt=1)
    try:
        idx = pcmx.update_on_gpu(-1, "nu", 2e-2)
        assert idx == 0
        pcmx.time_step(2, 0.0, c["dt"])
        out = np.zeros_like(f)
        pcmx.copy_farray_from_gpu(out)
    finally:
        pcmx.finalize_gpu()
    op = oracle_params(gp, c["dt"])
    op.nu = 2e-2
    o1, _, _, _ = O.advance(f0, op, 2)
    assert np.all(rel_linf(pcmx.interior_of(out, 4), o1) <= FP64_1STEP)


def _steps(f0, gp, nsteps, dt, timing, calls=1, change=None, forcing=None):
    f = pcmx.farray_from_interior(f0)
    pcmx.initialize_gpu(f, gp, nt=1)
    try:
        if forcing:
            pcmx.set_forcing(**forcing)
        pcmx.set_timing(timing)
        t = 0.0
        for k in range(calls):
            if change and k == 1:
                change()
            t, _ = pcmx.time_step(nsteps, t, dt, ldiag=True)
        red = pcmx.get_gpu_reduced_vars()
        out = np.zeros_like(f)
        pcmx.copy_farray_from_gpu(out)
    finally:
        pcmx.finalize_gpu()
    return pcmx.interior_of(out, gp.mvar).copy(), t, red


@pytest.mark.parametrize("name,nsteps", [("hydro32", 6), ("mhd32", 5), ("mhd_ragged", 4)])
def test_graph_replay_matches_plain_launches(name, nsteps):
    """Fixed steps replay captured pairs of steps (kernel timing off); the
    state, t and diagnostics equal those of plain launches bitwise, across
    repeated replays, an odd step count and a recapture after a parameter
    change and a new dt."""
    c, gp, f0 = setup_case(name)
    announce(f"test_graph_replay_matches_plain_launches[{name}] {nsteps} steps x 3 calls")
    change = lambda: pcmx.update_on_gpu(-1, "nu", 2 * c["nu"])  # noqa: E731
    a = _steps(f0, gp, nsteps, c["dt"], timing=False, calls=3, change=change)
    b = _steps(f0, gp, nsteps, c["dt"], timing=True, calls=3, change=change)
    np.testing.assert_array_equal(a[0], b[0])
    assert a[1] == b[1]
    np.testing.assert_array_equal(a[2], b[2])
    # a different dt recaptures
    a2 = _steps(f0, gp, nsteps, 0.5 * c["dt"], timing=False)
    b2 = _steps(f0, gp, nsteps, 0.5 * c["dt"], timing=True)
    np.testing.assert_array_equal(a2[0], b2[0])
