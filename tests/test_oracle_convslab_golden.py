"""Golden vector of the samples/conv-slab physics: the reference's own
double-precision run of the sample (samples/conv-slab/reference.out.double,
32^3, start.in / run.in: piecewise-polytropic stratification with an
isothermal top, 'up-down' u_z, gravity, Omega = 0.1, K-profile conduction,
top cooling, bcz s, s, a, a2, c1:cT, upwinded lnrho and ss, Courant steps)
reproduced by the oracle from the same initial state: t, dt, urms, umax,
rhom and ssm at it = 0, 10, 20, 30, 40, to the print precision of the file
(urms and umax carry 14 decimals, about 12 significant digits).  The rows
are committed as data in tests/golden/conv_slab_reference_double.json."""
import json
import os

import numpy as np
import pytest

import oracle as O
import pcmx_ics

HERE = os.path.dirname(__file__)
RTOL = {"t": 1e-13, "dt": 1e-13, "urms": 2e-12, "umax": 2e-12, "rhom": 1e-13, "ssm": 1e-13}


# (fixture, cp, hcond0): samples/conv-slab and samples/conv-slab_cp_2 (cp = 2,
# hcond0 doubled; its file prints mean(ss)/cp, the same physical run)
VARIANTS = {"cp1": ("conv_slab_reference_double.json", 1.0, 8e-3),
            "cp2": ("conv_slab_cp2_reference_double.json", 2.0, 1.6e-2)}


def golden_rows(variant="cp1"):
    with open(os.path.join(HERE, "golden", VARIANTS[variant][0])) as fh:
        return json.load(fh)["rows"]


def conv_slab_layers(mod, variant="cp1"):
    _, cp, hcond0 = VARIANTS[variant]
    return mod.layer_defaults(1.6666666, cp, -1.0, "c1:cT", hcond0=hcond0, mpoly0=1.0, mpoly1=3.0, mpoly2=0.0,
                              widthss=0.05, z1=0.0, z2=1.0, cool=15.0, wcool=0.2, cs2cool=0.333333)


CONV_SLAB_KW = dict(lentropy=True, nu=4e-3, gamma=1.6666666, cs0=0.57735, cdt=0.4, cdtv=0.3, L=(1.0, 1.0, 2.0),
                    gravz=-1.0, bcz=["s", "s", "a", "a2", "c1:cT"], omega=0.1, z0=-0.68, lupw_lnrho=True,
                    lupw_ss=True)


def check_row(ours, ref):
    for k, tol in RTOL.items():
        assert ours[k] == pytest.approx(ref[k], rel=tol, abs=1e-300), (ref["it"], k, ours[k], ref[k])


def means_after_bcs(interior, p):
    """rhom and mean(ss)/cp as the reference's diagnostics see them: after
    the boundary conditions of the pde call ('cT' sets ss on the top plane)."""
    g = O.with_ghosts(interior)
    O.boundconds(g, p)
    gi = O.interior(g)
    return float(np.exp(gi[3]).mean()), float(gi[4].mean()) / p.cp


@pytest.mark.parametrize("variant", ["cp1", "cp2"])
def test_oracle_reproduces_the_reference_run(variant):
    rows = golden_rows(variant)
    cp = VARIANTS[variant][1]
    p = O.make_params((32, 32, 32), dt=0.0, cp=cp, layers=conv_slab_layers(O, variant), **CONV_SLAB_KW)
    f = pcmx_ics.conv_slab_state(32, cp=cp)
    t = 0.0
    for it in range(41):
        f1, t1, dt, ser = O.advance(f, p, 1, t=t, diag=True)
        if it % 10 == 0:
            rhom, ssm = means_after_bcs(f, p)
            check_row(dict(t=t, dt=dt, urms=ser[0, 0], umax=ser[0, 1], rhom=rhom, ssm=ssm), rows[it // 10])
        f, t = f1, t1


def test_initial_state_matches_the_first_row():
    r = golden_rows()[0]
    f = pcmx_ics.conv_slab_state(32)
    u2 = (f[:3] ** 2).sum(0)
    p = O.make_params((32, 32, 32), dt=0.0, layers=conv_slab_layers(O), **CONV_SLAB_KW)
    rhom, ssm = means_after_bcs(f, p)
    assert np.sqrt(u2.mean()) == pytest.approx(r["urms"], rel=2e-12)
    assert np.sqrt(u2.max()) == pytest.approx(r["umax"], rel=2e-12)
    assert rhom == pytest.approx(r["rhom"], rel=1e-13) and ssm == pytest.approx(r["ssm"], rel=1e-13)
