"""Parity of every RHS kernel variant with the CPU oracle on grids that the
z-marching kernels accept (nx a multiple of 32, ny a multiple of 8).

Kernel 1 evaluates each stencil in the reference's operand order; kernels 2
and 3 march z planes through an LDS ring and split the xz/yz mixed
derivatives into a forward and a carried half (same terms, different
rounding).  Tolerances are the north-star ones: FP64 relative L-inf per
variable <= 1e-12 after one RK3 step, <= 1e-9 after many; the mixed form
<= 1e-5 after 100 steps."""
import math

import numpy as np
import pytest

import oracle as O
import pcmx
import pcmx_ics

pytestmark = pytest.mark.gpu

KERNELS = [1, 2, 3]

CASES = {
    # isothermal hydro, several z chunks per column
    "hydro_z": dict(n=(32, 16, 48), kind="noise", mag=False, ent=False, nu=5e-3, dt=1e-3),
    # polytropic hydro (noentropy with gamma /= 1)
    "poly_z": dict(n=(32, 8, 16), kind="noise", mag=False, ent=False, nu=5e-3, gamma=5 / 3, dt=1e-3),
    # isothermal MHD, two x tiles, two y tiles, several z chunks
    "mhd_z": dict(n=(64, 16, 32), kind="shells", mag=True, ent=False, nu=5e-3, eta=4e-3, dt=2e-3,
                  L=(2 * math.pi, 3.0, 4.0)),
    # hydro + entropy, fixed dt
    "hydross_z": dict(n=(32, 8, 24), kind="noise", mag=False, ent=True, nu=2e-3, chi=3e-3, gamma=5 / 3,
                      dt=1e-3),
    # MHD + entropy, Courant-limited dt
    "mhdss_z": dict(n=(32, 16, 24), kind="noise", mag=True, ent=True, nu=2e-3, eta=2e-3, chi=2e-3,
                    gamma=5 / 3, dt=0.0, ldt=True, cdt=0.4),
}


def rel_linf(a, b):
    out = []
    for v in range(a.shape[0]):
        den = np.abs(b[v]).max()
        num = np.abs(a[v] - b[v]).max()
        out.append(num / den if den > 0 else num)
    return np.array(out)


def setup(name, kernel, precision=pcmx.PREC_FP64, seed=11):
    c = dict(eta=0.0, chi=0.0, gamma=1.0, ldt=False, cdt=0.9, L=2 * math.pi)
    c.update(CASES[name])
    gp = pcmx.make_params(c["n"], lmagnetic=c["mag"], lentropy=c["ent"], nu=c["nu"], eta=c["eta"],
                          chi=c["chi"], gamma=c["gamma"], ldt=c["ldt"], cdt=c["cdt"], precision=precision,
                          kernel=kernel, L=c["L"])
    op = O.make_params(c["n"], lmagnetic=c["mag"], lentropy=c["ent"], nu=c["nu"], eta=c["eta"], chi=c["chi"],
                       gamma=c["gamma"], cdt=c["cdt"], dt=0.0 if c["ldt"] else c["dt"], L=c["L"])
    f0 = pcmx_ics.config_state(c["kind"], c["n"], seed, lmagnetic=c["mag"], lentropy=c["ent"])
    if c["ent"]:
        f0[4] += 0.1  # a non-zero mean entropy, so the mixed form's offsets matter
    return c, gp, op, f0


def announce(msg):
    print(msg, flush=True)


SMALL = list(CASES)


@pytest.mark.parametrize("kernel", KERNELS)
@pytest.mark.parametrize("name", SMALL)
def test_one_step(name, kernel):
    c, gp, op, f0 = setup(name, kernel)
    announce(f"test_one_step[{name}, kernel {kernel}]: n={c['n']}")
    g1, t, dt = pcmx.advance(f0, gp, 1, c["dt"])
    o1, to, dto, _ = O.advance(f0, op, 1)
    err = rel_linf(g1, o1)
    announce(f"  rel L-inf per variable: {err}")
    assert np.all(err <= 1e-12), err
    assert abs(t - to) <= 1e-15 * max(1.0, abs(to))
    if c["ldt"]:
        assert abs(dt - dto) <= 1e-13 * dto


@pytest.mark.parametrize("name", SMALL)
def test_twenty_steps_split_kernel(name):
    c, gp, op, f0 = setup(name, 3)
    announce(f"test_twenty_steps_split_kernel[{name}]")
    g1, _, _ = pcmx.advance(f0, gp, 20, c["dt"])
    o1, _, _, _ = O.advance(f0, op, 20)
    err = rel_linf(g1, o1)
    announce(f"  rel L-inf per variable: {err}")
    assert np.all(err <= 1e-9), err


@pytest.mark.parametrize("kernel", [2, 3])
def test_diagnostics_and_dt1(kernel):
    c, gp, op, f0 = setup("mhdss_z", kernel)
    announce(f"test_diagnostics_and_dt1[kernel {kernel}]")
    f = pcmx.farray_from_interior(f0)
    pcmx.initialize_gpu(f, gp, nt=1)
    try:
        dt1 = pcmx.gpu_prepare_for_first_substep(0.0)
        dt = pcmx.rhs_gpu(1, 0.0, 0.0)
        red = pcmx.get_gpu_reduced_vars()
    finally:
        pcmx.finalize_gpu()
    _, dt1o, (urms, umax, brms) = O.rhs(f0, op, want_dt=True)
    assert abs(dt1 - dt1o) <= 1e-13 * dt1o
    assert abs(red[3] - dt1o) <= 1e-13 * dt1o
    assert abs(dt - 1.0 / dt1o) <= 1e-13 / dt1o
    assert abs(red[0] - urms) <= 1e-13 * urms
    assert abs(red[1] - umax) <= 1e-13 * umax
    assert abs(red[2] - brms) <= 1e-13 * brms
    np.testing.assert_allclose(red[5:13], f0.reshape(8, -1).mean(axis=1), rtol=1e-13, atol=1e-16)


@pytest.mark.parametrize("name", ["mhd_z", "mhdss_z"])
def test_mixed_form_split_kernel(name):
    c, gp, op, f0 = setup(name, 3, precision=pcmx.PREC_MIXED)
    steps = 100 if not c["ldt"] else 60
    announce(f"test_mixed_form_split_kernel[{name}]: {steps} steps")
    g1, _, _ = pcmx.advance(f0, gp, steps, c["dt"])
    o1, _, _, _ = O.advance(f0, op, steps)
    err = rel_linf(g1, o1)
    announce(f"  mixed rel L-inf per variable: {err}")
    assert np.all(err <= 1e-5), err


def test_split_kernel_rejects_untiled_grid():
    gp = pcmx.make_params((24, 16, 16), kernel=3)
    with pytest.raises(pcmx.PcmxError, match="z-marching kernel needs"):
        pcmx.initialize_gpu(None, gp, nt=0)
    pcmx.finalize_gpu()


@pytest.mark.parametrize("name", ["mhd_z", "mhdss_z"])
def test_state_after_first_substep(name):
    """Kernel 3 stores only w1 on substep 1 and forms f1 = f0 + dt*beta1*w1 on
    the fly in substep 2; a copy-back in between must still return f1, as the
    direct kernel (which stores f1) does."""
    outs = []
    for kernel in (1, 3):
        c, gp, op, f0 = setup(name, kernel)
        f = pcmx.farray_from_interior(f0)
        pcmx.initialize_gpu(f, gp, nt=1)
        try:
            dt = pcmx.rhs_gpu(1, 0.0, c["dt"])
            out = np.zeros_like(f)
            pcmx.copy_farray_from_gpu(out)
            dt = pcmx.rhs_gpu(2, 0.0, dt)
            dt = pcmx.rhs_gpu(3, 0.0, dt)
            out3 = np.zeros_like(f)
            pcmx.copy_farray_from_gpu(out3)
        finally:
            pcmx.finalize_gpu()
        outs.append((pcmx.interior_of(out, gp.mvar), pcmx.interior_of(out3, gp.mvar)))
    announce(f"test_state_after_first_substep[{name}]")
    assert np.all(rel_linf(outs[1][0], outs[0][0]) <= 1e-13)
    assert np.all(rel_linf(outs[1][1], outs[0][1]) <= 1e-12)


# Grids with more blocks than the GPU runs at once (kernel 3 runs one block per
# CU with a magnetic field): later blocks read halos of tiles whose blocks have
# already finished, so a block that overwrote data its neighbours still read
# would show here.
WIDE = {
    "mhd_wide": dict(n=(512, 64, 32), kind="noise", mag=True, ent=False, nu=5e-3, eta=4e-3, dt=2e-3),
    "mhdss_wide": dict(n=(512, 64, 24), kind="noise", mag=True, ent=True, nu=2e-3, eta=2e-3, chi=2e-3,
                       gamma=5 / 3, dt=0.0, ldt=True, cdt=0.4),
}
CASES.update(WIDE)


@pytest.mark.parametrize("name", list(WIDE))
def test_multi_round_grid(name):
    c, gp3, op, f0 = setup(name, 3)
    _, gp1, _, _ = setup(name, 1)
    announce(f"test_multi_round_grid[{name}]: n={c['n']}")
    g3, t3, dt3 = pcmx.advance(f0, gp3, 3, c["dt"])
    g1, t1, dt1 = pcmx.advance(f0, gp1, 3, c["dt"])
    err = rel_linf(g3, g1)
    announce(f"  kernel 3 vs direct kernel, 3 steps: {err}")
    assert np.all(err <= 1e-12), err
    assert t3 == t1 or abs(t3 - t1) <= 1e-15 * abs(t1)
    o1, _, _, _ = O.advance(f0, op, 1)
    h1, _, _ = pcmx.advance(f0, gp3, 1, c["dt"])
    err1 = rel_linf(h1, o1)
    announce(f"  kernel 3 vs oracle, 1 step: {err1}")
    assert np.all(err1 <= 1e-12), err1


@pytest.mark.parametrize("name", ["hydro_z", "mhd_z", "mhdss_z"])
def test_tuned_plan_is_bitwise_identical(name):
    """pcmx_tune only changes the tile shape and z-chunk length, which do not
    change any cell's arithmetic, so results are bitwise equal."""
    c, gp, op, f0 = setup(name, 3)
    announce(f"test_tuned_plan_is_bitwise_identical[{name}]")
    ref, _, _ = pcmx.advance(f0, gp, 3, c["dt"])
    f = pcmx.farray_from_interior(f0)
    pcmx.initialize_gpu(f, gp, nt=1)
    try:
        ms, tile, lz = pcmx.tune()
        announce(f"  tuned: {ms:.4f} ms per launch, tile width {tile}, {lz}-plane chunks")
        pcmx.time_step(3, 0.0, c["dt"])
        out = np.zeros_like(f)
        pcmx.copy_farray_from_gpu(out)
    finally:
        pcmx.finalize_gpu()
    np.testing.assert_array_equal(pcmx.interior_of(out, gp.mvar), ref)
