"""Parity of the samples/conv-slab layer physics on the HIP path (direct
kernel, and the split kernel's entropy variant with lnrho + ss upwinding)
with the CPU oracle, through the libpcmx C ABI: Coriolis force
(coriolis_cartesian, hydro.f90:6339-6413), K-profile heat conduction
(calc_heatcond, entropy.f90:5913-6151), heating and cooling
(get_heat_cool_gravz, :6761-6828) and the entropy boundary rules 'c1'
(bc_ss_flux) and 'cT' (bc_ss_temp_z), with gravity, 's'/'a'/'a2' for the
other variables and upwinded lnrho and ss, as samples/conv-slab runs them.
FP64 relative L-inf per variable <= 1e-12 after one step and <= 1e-9 after
20 Courant steps (dt itself to 1e-12); the mixed form <= 1e-5 after 100."""
import math

import numpy as np
import pytest

import oracle as O
import pcmx
import pcmx_ics
from test_gpu_parity import FP64_1STEP, FP64_100STEP, MIXED_100STEP, announce, rel_linf

pytestmark = pytest.mark.gpu

GAMMA = 1.6666666
CS0 = 0.57735
L = (1.0, 1.0, 2.0)
Z0 = -0.68
BCZ = ["s", "s", "a", "a2", "c1:cT"]


def layers(mod, bcz_ss="c1:cT", **kw):
    base = dict(hcond0=8e-3, mpoly0=1.0, mpoly1=3.0, mpoly2=0.0, widthss=0.05, z1=0.0, z2=1.0,
                wheat=0.1, cool=15.0, wcool=0.2, cs2cool=0.333333)
    base.update(kw)
    return mod.layer_defaults(GAMMA, 1.0, -1.0, bcz_ss, **base)


def conv_state(n, seed):
    f0 = pcmx_ics.config_state("noise", n, seed, lentropy=True)
    nz = n[2]
    z = Z0 + L[2] / (nz - 1) * np.arange(nz)[:, None, None]
    f0[:3] *= 0.5
    f0[3] += -0.8 * (z - Z0)                  # stratified density
    f0[4] = 0.05 * f0[4] - 0.1 * (z - Z0)     # entropy decreasing upwards
    return f0


def both(n, bcz=BCZ, omega=0.1, theta=0.0, lay_kw=None, ldt=True, dt=2e-4, steps=1, seed=3,
         precision=pcmx.PREC_FP64, lupw=(False, True, True), gravz=-1.0, kernel=0):
    lay_kw = dict(lay_kw or {})
    ss_rule = bcz[4] if len(bcz) > 4 else "s"
    kw = dict(lentropy=True, nu=4e-3, gamma=GAMMA, cs0=CS0, cdt=0.4, cdtv=0.3, L=L, gravz=gravz, bcz=bcz,
              omega=omega, theta=theta, z0=Z0, lupw_uu=lupw[0], lupw_lnrho=lupw[1], lupw_ss=lupw[2])
    gp = pcmx.make_params(n, ldt=ldt, kernel=kernel, precision=precision, layers=layers(pcmx, ss_rule, **lay_kw),
                          **kw)
    op = O.make_params(n, dt=0.0 if ldt else dt, layers=layers(O, ss_rule, **lay_kw), **kw)
    f0 = conv_state(n, seed)
    g, t, gdt = pcmx.advance(f0, gp, steps, dt)
    o, to, odt, _ = O.advance(f0, op, steps)
    return f0, g, o, (t, gdt, to, odt)


@pytest.mark.parametrize("kernel", [1, 3])
def test_conv_slab_one_and_twenty_courant_steps(kernel):
    """kernel 3: the split kernel's entropy variant with lnrho + ss upwinding
    and the layer terms (32x8 tiles); kernel 1 the direct kernel."""
    n = (32, 16, 24)
    for steps, tol in ((1, FP64_1STEP), (20, FP64_100STEP)):
        announce(f"test_conv_slab_one_and_twenty_courant_steps[kernel {kernel}] n={n} steps={steps}")
        f0, g, o, (t, gdt, to, odt) = both(n, steps=steps, kernel=kernel)
        err = rel_linf(g, o)
        announce(f"  rel L-inf per variable: {err}  dt gpu {gdt!r} oracle {odt!r}")
        assert np.all(err <= tol), err
        assert gdt == pytest.approx(odt, rel=FP64_1STEP) and t == pytest.approx(to, rel=1e-12)


@pytest.mark.parametrize("kernel", [1, 3])
@pytest.mark.parametrize("piece", ["rotation60", "kprof", "cool_step", "heating", "c1_top_ct_bottom"])
def test_each_piece_one_and_two_steps(piece, kernel):
    n = (32, 8, 16)
    announce(f"test_each_piece_one_and_two_steps[{piece}, kernel {kernel}] n={n}")
    kw = dict(omega=0.0, lay_kw=dict(hcond0=0.0, cool=0.0, cs2cool=0.0, kprof=False), bcz=["s", "s", "a", "a2", "a2"])
    if piece == "rotation60":
        kw.update(omega=0.23, theta=60.0)
    elif piece == "kprof":
        kw["lay_kw"] = dict(cool=0.0, cs2cool=0.0, widthss=0.2)
    elif piece == "cool_step":
        # fixed steps: the Courant step ignores the cooling rate, as the
        # reference's does (lthdiff_Hmax=F), and without conduction it is too
        # long for cool=15
        kw.update(ldt=False, dt=2e-4)
        kw["lay_kw"] = dict(hcond0=0.0, kprof=False, cooling_profile="step", wcool=0.1)
    elif piece == "heating":
        kw.update(ldt=False, dt=2e-4)
        kw["lay_kw"] = dict(hcond0=0.0, kprof=False, cool=0.0, luminosity=0.3, wheat=0.15, cs2cool=0.4)
    else:
        kw.update(bcz=["s", "s", "a", "a2", "cT:c1"], omega=0.1)
        kw["lay_kw"] = dict(cs2bot=0.5, ftop=0.004)
    for steps, tol in ((1, FP64_1STEP), (2, 10 * FP64_1STEP)):
        f0, g, o, _ = both(n, steps=steps, kernel=kernel, **kw)
        err = rel_linf(g, o)
        announce(f"  {steps} steps: rel L-inf per variable: {err}")
        assert np.all(err <= tol), err
    # the piece changes the answer
    base = O.advance(f0, O.make_params(n, lentropy=True, nu=4e-3, gamma=GAMMA, cs0=CS0, cdt=0.4, cdtv=0.3, L=L,
                                       gravz=-1.0, bcz=["s", "s", "a", "a2", "a2"], z0=Z0, lupw_lnrho=True,
                                       lupw_ss=True, dt=kw.get("dt", 0.0) if not kw.get("ldt", True) else 0.0),
                     2)[0]
    assert rel_linf(o, base).max() > 1e-9


@pytest.mark.parametrize("kernel", [1, 3])
def test_magnetic_conv_slab(kernel):
    """The conv-slab physics with a magnetic field (bcz a, a, s for A):
    Lorentz force and ohmic heating beside the layer terms."""
    n = (32, 16, 24)
    bcz = BCZ + ["a", "a", "s"]
    kw = dict(lentropy=True, lmagnetic=True, nu=4e-3, eta=3e-3, gamma=GAMMA, cs0=CS0, cdt=0.4, cdtv=0.3, L=L,
              gravz=-1.0, bcz=bcz, omega=0.1, z0=Z0, lupw_lnrho=True, lupw_ss=True)
    f0 = pcmx_ics.config_state("shells", n, 9, lmagnetic=True, lentropy=True)
    ref = conv_state(n, 9)
    f0[:5] = ref
    f0[5:] *= 0.3
    for steps, tol in ((1, FP64_1STEP), (10, FP64_100STEP)):
        announce(f"test_magnetic_conv_slab[kernel {kernel}] n={n} steps={steps}")
        gp = pcmx.make_params(n, ldt=True, kernel=kernel, layers=layers(pcmx), **kw)
        op = O.make_params(n, dt=0.0, layers=layers(O), **kw)
        g, t, gdt = pcmx.advance(f0, gp, steps, 0.0)
        o, to, odt, _ = O.advance(f0, op, steps)
        err = rel_linf(g, o)
        announce(f"  rel L-inf per variable: {err}")
        assert np.all(err <= tol), err
        assert gdt == pytest.approx(odt, rel=FP64_1STEP)


@pytest.mark.parametrize("kernel", [1, 3])
def test_entropy_rules_read_a_zeroed_lnrho_boundary(kernel):
    """lnrho with 'a' (its boundary value zeroed before the entropy rules
    read it, as the variable loop of boundconds_z orders them) under
    'c1:cT' and 'cT:c1' for ss: one step."""
    n = (32, 8, 16)
    for ss_rule, lay_kw in (("c1:cT", {}), ("cT:c1", dict(cs2bot=0.5, ftop=0.004))):
        announce(f"test_entropy_rules_read_a_zeroed_lnrho_boundary[kernel {kernel}] ss {ss_rule}")
        f0, g, o, _ = both(n, bcz=["s", "s", "a", "a", ss_rule], kernel=kernel, lay_kw=lay_kw)
        err = rel_linf(g, o)
        announce(f"  rel L-inf per variable: {err}")
        assert np.all(err <= FP64_1STEP), err


def test_isothermal_rotation_with_split_kernel_masks():
    """Rotation without entropy on the direct kernel, with uu + lnrho
    upwinding (the published upwinding set)."""
    n = (32, 16, 16)
    announce(f"test_isothermal_rotation_with_split_kernel_masks n={n}")
    gp = pcmx.make_params(n, nu=4e-3, ldt=True, cdt=0.4, kernel=0, omega=0.5, lupw_uu=True, lupw_lnrho=True)
    op = O.make_params(n, nu=4e-3, cdt=0.4, dt=0.0, omega=0.5, lupw_uu=True, lupw_lnrho=True)
    f0 = pcmx_ics.config_state("noise", n, 7)
    g, t, gdt = pcmx.advance(f0, gp, 3, 0.0)
    o, to, odt, _ = O.advance(f0, op, 3)
    err = rel_linf(g, o)
    announce(f"  rel L-inf per variable: {err}")
    assert np.all(err <= FP64_1STEP * 10), err


@pytest.mark.parametrize("kernel", [1, 3])
def test_mixed_form_hundred_steps(kernel):
    n = (32, 16, 24)
    announce(f"test_mixed_form_hundred_steps[kernel {kernel}] n={n}")
    f0, g, o, _ = both(n, steps=100, ldt=False, dt=1e-4, precision=pcmx.PREC_MIXED, kernel=kernel)
    err = rel_linf(g, o)
    announce(f"  rel L-inf per variable: {err}")
    assert np.all(err <= MIXED_100STEP), err


def test_copied_ghosts_follow_c1_and_ct():
    n = (32, 16, 16)
    announce("test_copied_ghosts_follow_c1_and_ct")
    gp = pcmx.make_params(n, lentropy=True, nu=4e-3, gamma=GAMMA, cs0=CS0, L=L, gravz=-1.0, bcz=BCZ, z0=Z0,
                          omega=0.1, layers=layers(pcmx), kernel=0)
    f0 = conv_state(n, 5)
    f = pcmx.farray_from_interior(f0)
    pcmx.initialize_gpu(f, gp, nt=1)
    try:
        pcmx.time_step(2, 0.0, 2e-4)
        out = np.zeros_like(f)
        pcmx.copy_farray_from_gpu(out)
    finally:
        pcmx.finalize_gpu()
    ref = O.with_ghosts(pcmx.interior_of(out, 5).copy())
    O.boundconds(ref, O.make_params(n, lentropy=True, gamma=GAMMA, cs0=CS0, L=L, bcz=BCZ, z0=Z0,
                                    layers=layers(O)))
    g = 3
    for plane in list(range(0, g)) + list(range(g + n[2], 2 * g + n[2])):
        np.testing.assert_array_equal(out[:4, plane], ref[:4, plane])
        np.testing.assert_allclose(out[4, plane], ref[4, plane], rtol=1e-14, atol=1e-15)


def test_rejected_layer_configurations():
    announce("test_rejected_layer_configurations")
    n = (32, 8, 16)
    kw = dict(lentropy=True, gamma=GAMMA, L=L, bcz=BCZ, z0=Z0)
    with pytest.raises(pcmx.PcmxError, match="kernel 1"):    # the split kernel needs lnrho + ss upwinding
        pcmx.initialize_gpu(None, pcmx.make_params(n, kernel=3, layers=layers(pcmx), **kw), nt=0)
    with pytest.raises(pcmx.PcmxError, match="32x8 tiles"):  # and 32x8 tiles
        pcmx.initialize_gpu(None, pcmx.make_params((16, 16, 16), kernel=3, layers=layers(pcmx), lupw_lnrho=True,
                                                   lupw_ss=True, **kw), nt=0)
    no_k = layers(pcmx)
    no_k["lheatc_kprof"] = 0
    with pytest.raises(pcmx.PcmxError, match="'c1' needs"):
        pcmx.initialize_gpu(None, pcmx.make_params(n, kernel=0, layers=no_k, **kw), nt=0)
    no_t = layers(pcmx, cs2top=0.0)
    with pytest.raises(pcmx.PcmxError, match="'cT' needs cs2top"):
        pcmx.initialize_gpu(None, pcmx.make_params(n, kernel=0, layers=no_t, **kw), nt=0)
    with pytest.raises(pcmx.PcmxError, match="need the entropy"):
        pcmx.initialize_gpu(None, pcmx.make_params(n, kernel=0, L=L, z0=Z0, layers=layers(pcmx, "s"),
                                                   bcz=["s", "s", "a", "a2"]), nt=0)
    with pytest.raises(pcmx.PcmxError, match="entropy slot"):
        pcmx.initialize_gpu(None, pcmx.make_params(n, kernel=0, L=L, z0=Z0, bcz=["s", "s", "a", "c1"]), nt=0)
