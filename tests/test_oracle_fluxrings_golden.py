"""Golden vector of isothermal MHD: the reference's own double-precision
run of samples/interlocked-fluxrings (reference.out.double, 32^3 periodic
box of side 2 pi, nu = eta = 3e-3, Courant steps with cdt = 0.9, two
interlocked flux rings in A, u = 0, lnrho = 0) reproduced by the oracle from
the same initial state: t, urms, umax and brms at it = 0, 2, 4, 6, 8, to the
print precision of the file (12-14 significant digits).  The sample's
radiation_ray module is configured for visualisation only (source and
opacity from B^2, no energy equation) and adds no term here.  The rows are
committed as data in tests/golden/fluxrings_reference_double.json."""
import json
import os

import pytest

import oracle as O
import pcmx_ics

HERE = os.path.dirname(__file__)
RTOL = {"t": 1e-12, "urms": 1e-12, "umax": 2e-12, "brms": 1e-12}


def golden_rows():
    with open(os.path.join(HERE, "golden", "fluxrings_reference_double.json")) as fh:
        return json.load(fh)["rows"]


FLUXRINGS_KW = dict(lmagnetic=True, nu=3e-3, eta=3e-3, cdt=0.9)


def check_row(ours, ref):
    for k, tol in RTOL.items():
        assert ours[k] == pytest.approx(ref[k], rel=tol, abs=1e-300), (ref["it"], k, ours[k], ref[k])


def test_oracle_reproduces_the_reference_run():
    rows = golden_rows()
    p = O.make_params(32, dt=0.0, **FLUXRINGS_KW)
    f = pcmx_ics.fluxrings_state(32)
    t = 0.0
    for it in range(9):
        f1, t1, dt, ser = O.advance(f, p, 1, t=t, diag=True)
        if it % 2 == 0:
            check_row(dict(t=t, urms=ser[0, 0], umax=ser[0, 1], brms=ser[0, 2]), rows[it // 2])
        f, t = f1, t1
