"""Parity of upwinding (lupw_uu, lupw_lnrho) and helical forcing on the HIP
path with the CPU oracle, through the libpcmx C ABI, at the same bars as
the plain path: FP64 relative L-inf per variable <= 1e-12 after one step and
<= 1e-9 after many, the mixed form <= 1e-5.  Each case also checks that the
option changes the oracle's answer by far more than the tolerance, so the
comparison exercises it."""
import math

import numpy as np
import pytest

import oracle as O
import pcmx
import pcmx_forcing as PF
import pcmx_ics
from test_gpu_parity import (FP64_1STEP, FP64_100STEP, MIXED_100STEP, announce, oracle_params, rel_linf,
                             setup_case)

pytestmark = pytest.mark.gpu

UPW = {"both": (True, True), "uu": (True, False), "lnrho": (False, True)}


def run_pair(name, steps, kernel, flags, precision=pcmx.PREC_FP64, forcing=None, ofo=None):
    c, gp, f0 = setup_case(name, precision=precision, kernel=kernel, lupw_uu=flags[0], lupw_lnrho=flags[1])
    g, t, dt = pcmx.advance(f0, gp, steps, c["dt"], forcing=forcing)
    op = oracle_params(gp, c["dt"])
    o, to, dto, _ = O.advance(f0, op, steps, forcing=ofo)
    return c, gp, f0, g, o, (t, dt, to, dto)


def assert_exercised(f0, gp, c, steps, o, ofo=None, plain_flags=(False, False)):
    op = oracle_params(gp, c["dt"])
    op.lupw_uu, op.lupw_lnrho = plain_flags
    plain, _, _, _ = O.advance(f0, op, steps)
    assert rel_linf(o, plain).max() > 1e3 * FP64_1STEP


@pytest.mark.parametrize("name,kernel", [("hydro32", 1), ("hydro32", 3), ("mhd32", 1), ("mhd32", 3),
                                         ("mhd_ragged", 1)])
def test_upwinding_one_step(name, kernel):
    announce(f"test_upwinding_one_step[{name}, kernel {kernel}]")
    c, gp, f0, g, o, _ = run_pair(name, 1, kernel, UPW["both"])
    err = rel_linf(g, o)
    announce(f"  rel L-inf per variable: {err}")
    assert np.all(err <= FP64_1STEP), err
    assert_exercised(f0, gp, c, 1, o)


def test_upwinding_split_kernel_16x16_tiles():
    """nx = 48 is not a multiple of 32: the split kernel runs 16x16 tiles."""
    announce("test_upwinding_split_kernel_16x16_tiles: mhd 48x16x12, kernel 3")
    n, L, dt = (48, 16, 12), (2 * math.pi, 2.0, 1.5), 1e-3
    gp = pcmx.make_params(n, lmagnetic=True, nu=5e-3, eta=4e-3, kernel=3, L=L, lupw_uu=True, lupw_lnrho=True)
    f0 = pcmx_ics.config_state("shells", n, 5, lmagnetic=True)
    g, _, _ = pcmx.advance(f0, gp, 1, dt)
    o, _, _, _ = O.advance(f0, oracle_params(gp, dt), 1)
    err = rel_linf(g, o)
    announce(f"  rel L-inf per variable: {err}")
    assert np.all(err <= FP64_1STEP), err


@pytest.mark.parametrize("which", ["uu", "lnrho"])
def test_upwinding_single_flag_direct_kernel(which):
    announce(f"test_upwinding_single_flag_direct_kernel[{which}]: mhd32")
    c, gp, f0, g, o, _ = run_pair("mhd32", 1, 1, UPW[which])
    err = rel_linf(g, o)
    assert np.all(err <= FP64_1STEP), err
    assert_exercised(f0, gp, c, 1, o)


def test_upwinding_with_entropy_courant():
    announce("test_upwinding_with_entropy_courant: mhdss24, auto kernel")
    c, gp, f0, g, o, (t, dt, to, dto) = run_pair("mhdss24", 2, 0, UPW["both"])
    err = rel_linf(g, o)
    announce(f"  rel L-inf per variable: {err}")
    assert np.all(err <= FP64_1STEP * 10), err
    assert abs(dt - dto) <= 1e-13 * dto


@pytest.mark.parametrize("kernel", [1, 3])
def test_upwinding_many_steps(kernel):
    announce(f"test_upwinding_many_steps[kernel {kernel}]: mhd32, 60 steps")
    c, gp, f0, g, o, _ = run_pair("mhd32", 60, kernel, UPW["both"])
    err = rel_linf(g, o)
    announce(f"  rel L-inf per variable: {err}")
    assert np.all(err <= FP64_100STEP), err


def test_upwinding_mixed_form():
    announce("test_upwinding_mixed_form: mhd32, kernel 3, 100 steps")
    c, gp, f0, g, o, _ = run_pair("mhd32", 100, 3, UPW["both"], precision=pcmx.PREC_MIXED)
    err = rel_linf(g, o)
    announce(f"  mixed rel L-inf per variable: {err}")
    assert np.all(err <= MIXED_100STEP), err


def test_split_kernel_rejects_unsupported_upwinding():
    announce("test_split_kernel_rejects_unsupported_upwinding")
    for flags, ent in (((True, False), False), ((True, True), True)):
        gp = pcmx.make_params(16, lmagnetic=True, lentropy=ent, gamma=5 / 3 if ent else 1.0, kernel=3,
                              lupw_uu=flags[0], lupw_lnrho=flags[1])
        with pytest.raises(pcmx.PcmxError, match="upwind"):
            pcmx.initialize_gpu(None, gp, nt=0)


# ------------------------------------------------------------------ forcing
def forcing_pair(kvec, gp, phase=0.9, phi=2.2, force=0.07, dt=2e-3, kav=3.0, relhel=1.0, k1_ff=1.0,
                 target=0, scale_sqrt_dt=False, fact_dt=None):
    """Keyword arguments of pcmx.set_forcing and the oracle's Forcing for
    the same parameters (fact_dt: the step the oracle multiplies in when the
    library scales by sqrt(dt))."""
    pars = PF.helical_pars(kvec, phase, phi, force, dt, kav, relhel=relhel, scale_sqrt_dt=scale_sqrt_dt)
    xyz0 = (-math.pi, -1.5, -1.25)
    kw = dict(kk=pars["kk"], phase=pars["phase"], fact=pars["fact"], coef1=pars["coef1"], coef2=pars["coef2"],
              fda=pars["fda"], k1_ff=k1_ff, xyz0=xyz0, target=target, scale_sqrt_dt=scale_sqrt_dt)
    fact = pars["fact"] * math.sqrt(fact_dt) if scale_sqrt_dt else pars["fact"]
    ofo = O.make_forcing(pars["kk"], pars["phase"], fact, pars["coef1"], pars["coef2"], pars["fda"],
                         k1_ff=k1_ff, xyz0=xyz0, target=target)
    return kw, ofo


def test_forcing_on_a_state_at_rest():
    """f = 0: the step leaves f at 0 and the forcing is the whole answer."""
    announce("test_forcing_on_a_state_at_rest: 32x16x20 hydro, kernel 3")
    n = (32, 16, 20)
    gp = pcmx.make_params(n, nu=5e-3, kernel=3, L=(2 * math.pi, 3.0, 2.5))
    f0 = np.zeros((4, n[2], n[1], n[0]))
    kw, ofo = forcing_pair((1, 2, -2), gp, k1_ff=1.0)
    g, _, _ = pcmx.advance(f0, gp, 1, 1e-3, forcing=kw)
    ref = O.addforce(f0, oracle_params(gp, 1e-3), ofo)
    err = rel_linf(g[:3], ref[:3])
    announce(f"  rel L-inf: {err}")
    assert np.all(err <= 1e-14), err
    assert np.abs(g[3]).max() == 0.0


@pytest.mark.parametrize("kernel", [1, 3])
def test_forcing_mhd_steps(kernel):
    announce(f"test_forcing_mhd_steps[kernel {kernel}]: mhd32 forced on uu, 1 and 20 steps")
    c, gp, f0 = setup_case("mhd32", kernel=kernel)
    kw, ofo = forcing_pair((2, -1, 1), gp, force=0.2)
    for steps, tol in ((1, FP64_1STEP), (20, FP64_100STEP)):
        g, _, _ = pcmx.advance(f0, gp, steps, c["dt"], forcing=kw)
        o, _, _, _ = O.advance(f0, oracle_params(gp, c["dt"]), steps, forcing=ofo)
        err = rel_linf(g, o)
        announce(f"  {steps} steps: rel L-inf per variable: {err}")
        assert np.all(err <= tol), err
        plain, _, _, _ = O.advance(f0, oracle_params(gp, c["dt"]), steps)
        assert rel_linf(o, plain).max() > 1e3 * tol


def test_forcing_on_a_with_upwinding():
    """The physics of the reference's published timings: isothermal MHD with
    lupw_uu = lupw_lnrho = T and helical forcing of the vector potential
    (lmagnetic_forcing = T), split kernel."""
    announce("test_forcing_on_a_with_upwinding: mhd32, kernel 3, 20 steps")
    c, gp, f0 = setup_case("mhd32", kernel=3, lupw_uu=True, lupw_lnrho=True)
    kw, ofo = forcing_pair((1, 1, -2), gp, force=0.05, target=gp.iax)
    g, _, _ = pcmx.advance(f0, gp, 20, c["dt"], forcing=kw)
    o, _, _, _ = O.advance(f0, oracle_params(gp, c["dt"]), 20, forcing=ofo)
    err = rel_linf(g, o)
    announce(f"  rel L-inf per variable: {err}")
    assert np.all(err <= FP64_100STEP), err


def test_forcing_with_courant_dt_scales_by_sqrt_dt():
    announce("test_forcing_with_courant_dt_scales_by_sqrt_dt: mhdss24, 1 step")
    c, gp, f0 = setup_case("mhdss24")
    # the oracle's Courant dt of the first step, for the oracle's fact
    _, _, dto, _ = O.advance(f0, oracle_params(gp, 0.0), 1)
    kw, ofo = forcing_pair((0, 1, 1), gp, force=0.1, scale_sqrt_dt=True, fact_dt=dto)
    g, t, dt = pcmx.advance(f0, gp, 1, 0.0, forcing=kw)
    o, _, _, _ = O.advance(f0, oracle_params(gp, 0.0), 1, forcing=ofo)
    assert abs(dt - dto) <= 1e-13 * dto
    err = rel_linf(g, o)
    announce(f"  rel L-inf per variable: {err}")
    assert np.all(err <= FP64_1STEP), err


def test_forcing_per_substep_api_matches_time_step():
    announce("test_forcing_per_substep_api_matches_time_step: mhd32, 2 steps")
    c, gp, f0 = setup_case("mhd32", kernel=3)
    kw, _ = forcing_pair((2, 0, 1), gp)
    a, ta, _ = pcmx.advance(f0, gp, 2, c["dt"], per_substep=True, forcing=kw)
    b, tb, _ = pcmx.advance(f0, gp, 2, c["dt"], forcing=kw)
    np.testing.assert_array_equal(a, b)


def test_set_forcing_errors_and_switch_off():
    announce("test_set_forcing_errors_and_switch_off")
    c, gp, f0 = setup_case("hydro32")
    f = pcmx.farray_from_interior(f0)
    pcmx.initialize_gpu(f, gp, nt=1)
    try:
        kw, _ = forcing_pair((1, 0, 0), gp)
        with pytest.raises(pcmx.PcmxError, match="target"):
            pcmx.set_forcing(**dict(kw, target=1))
        with pytest.raises(pcmx.PcmxError, match="non-finite"):
            pcmx.set_forcing(**dict(kw, phase=float("nan")))
        pcmx.set_forcing(**kw)
        pcmx.clear_forcing()
        pcmx.time_step(1, 0.0, c["dt"])
        out = np.zeros_like(f)
        pcmx.copy_farray_from_gpu(out)
    finally:
        pcmx.finalize_gpu()
    o, _, _, _ = O.advance(f0, oracle_params(gp, c["dt"]), 1)
    assert np.all(rel_linf(pcmx.interior_of(out, 4), o) <= FP64_1STEP)


def test_graph_replay_with_upwinding_and_forcing():
    """Captured step pairs (kernel timing off) against plain launches, with
    upwinding and the fused forcing epilogue: bitwise equal."""
    from test_gpu_parity import _steps
    announce("test_graph_replay_with_upwinding_and_forcing: mhd32, kernel 3, 6 steps")
    c, gp, f0 = setup_case("mhd32", kernel=3, lupw_uu=True, lupw_lnrho=True)
    kw, _ = forcing_pair((1, -2, 1), gp, force=0.1, target=gp.iax)
    a = _steps(f0, gp, 6, c["dt"], timing=False, calls=2, forcing=kw)
    b = _steps(f0, gp, 6, c["dt"], timing=True, calls=2, forcing=kw)
    np.testing.assert_array_equal(a[0], b[0])
    assert a[1] == b[1]


@pytest.mark.parametrize("flags", [(False, False, True), (True, True, True), (False, True, True)])
def test_entropy_upwinding_direct_kernel(flags):
    """lupw_ss alone and with the other flags (the direct kernel; the split
    kernel rejects upwinding with entropy), Courant steps."""
    announce(f"test_entropy_upwinding_direct_kernel[{flags}]: mhdss24")
    c = {"dt": 0.0}
    gp = pcmx.make_params(24, lmagnetic=True, lentropy=True, nu=2e-3, eta=2e-3, chi=2e-3, gamma=5 / 3, ldt=True,
                          cdt=0.4, kernel=0, lupw_uu=flags[0], lupw_lnrho=flags[1], lupw_ss=flags[2])
    f0 = pcmx_ics.config_state("noise", 24, 7, lmagnetic=True, lentropy=True)
    g, _, dt = pcmx.advance(f0, gp, 2, 0.0)
    op = oracle_params(gp, 0.0)
    o, _, dto, _ = O.advance(f0, op, 2)
    err = rel_linf(g, o)
    announce(f"  rel L-inf per variable: {err}")
    assert np.all(err <= FP64_1STEP * 10), err
    assert abs(dt - dto) <= 1e-13 * dto
    op.lupw_ss = 0
    plain, _, _, _ = O.advance(f0, op, 2)
    assert rel_linf(o, plain)[4] > 1e3 * FP64_1STEP


def test_lupw_ss_needs_entropy():
    announce("test_lupw_ss_needs_entropy")
    with pytest.raises(pcmx.PcmxError, match="lupw_ss"):
        pcmx.initialize_gpu(None, pcmx.make_params(16, lupw_ss=True), nt=0)


@pytest.mark.parametrize("lmag", [True, False])
@pytest.mark.parametrize("flags", [(False, True, True), (True, True, True)])
def test_entropy_upwinding_split_kernel(flags, lmag):
    """Entropy with lnrho + ss or uu + lnrho + ss upwinding on the split
    kernel (32x8 tiles), Courant steps (the deferred first-substep update)."""
    n = (32, 16, 16)
    announce(f"test_entropy_upwinding_split_kernel[{flags}, mag={lmag}] n={n}")
    gp = pcmx.make_params(n, lmagnetic=lmag, lentropy=True, nu=2e-3, eta=2e-3, chi=2e-3, gamma=5 / 3, ldt=True,
                          cdt=0.4, kernel=3, lupw_uu=flags[0], lupw_lnrho=flags[1], lupw_ss=flags[2],
                          L=(2 * math.pi, math.pi, math.pi))
    f0 = pcmx_ics.config_state("shells" if lmag else "noise", n, 7, lmagnetic=lmag, lentropy=True)
    f0[4] += 0.01 * np.random.default_rng(3).standard_normal(f0[4].shape)
    g, _, dt = pcmx.advance(f0, gp, 3, 0.0)
    op = oracle_params(gp, 0.0)
    o, _, dto, _ = O.advance(f0, op, 3)
    err = rel_linf(g, o)
    announce(f"  rel L-inf per variable: {err}")
    assert np.all(err <= FP64_1STEP * 10), err
    assert abs(dt - dto) <= 1e-13 * dto
    op.lupw_uu = op.lupw_lnrho = op.lupw_ss = 0
    plain, _, _, _ = O.advance(f0, op, 3)
    assert rel_linf(o, plain)[4] > 1e3 * FP64_1STEP


def test_entropy_upwinding_split_kernel_mixed_form():
    n = (32, 16, 16)
    announce("test_entropy_upwinding_split_kernel_mixed_form: 60 steps")
    gp = pcmx.make_params(n, lmagnetic=True, lentropy=True, nu=2e-3, eta=2e-3, chi=2e-3, gamma=5 / 3, ldt=False,
                          kernel=3, lupw_uu=True, lupw_lnrho=True, lupw_ss=True, precision=pcmx.PREC_MIXED)
    f0 = pcmx_ics.config_state("shells", n, 7, lmagnetic=True, lentropy=True)
    f0[4] += 0.01 * np.random.default_rng(3).standard_normal(f0[4].shape)
    g, _, _ = pcmx.advance(f0, gp, 60, 1e-3)
    o, _, _, _ = O.advance(f0, oracle_params(gp, 1e-3), 60)
    err = rel_linf(g, o)
    announce(f"  mixed rel L-inf per variable: {err}")
    assert np.all(err <= MIXED_100STEP), err


def test_entropy_upwinding_16x16_grid_uses_the_direct_kernel():
    announce("test_entropy_upwinding_16x16_grid_uses_the_direct_kernel")
    kw = dict(lmagnetic=True, lentropy=True, gamma=5 / 3, lupw_lnrho=True, lupw_ss=True)
    with pytest.raises(pcmx.PcmxError, match="32x8"):
        pcmx.initialize_gpu(None, pcmx.make_params((48, 16, 16), kernel=3, **kw), nt=0)
    gp = pcmx.make_params((48, 16, 16), kernel=0, **kw)
    f0 = pcmx_ics.config_state("noise", (48, 16, 16), 2, lmagnetic=True, lentropy=True)
    g, _, _ = pcmx.advance(f0, gp, 1, 1e-3)
    o, _, _, _ = O.advance(f0, oracle_params(gp, 1e-3), 1)
    assert np.all(rel_linf(g, o) <= FP64_1STEP)
