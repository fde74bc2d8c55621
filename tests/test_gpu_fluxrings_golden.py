"""The reference's samples/interlocked-fluxrings run (reference.out.double,
isothermal MHD 32^3, Courant steps) on the HIP path through the C ABI, on
the split kernel (the automatic choice) and on the direct kernel: t, urms,
umax and brms at it = 0, 2, 4, 6, 8 against the reference's printed values,
with the oracle's golden tolerances."""
import pytest

import pcmx
import pcmx_ics
from test_gpu_parity import announce
from test_oracle_fluxrings_golden import FLUXRINGS_KW, check_row, golden_rows

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("kernel", [0, 1])
def test_fluxrings_sample_reproduces_the_reference_run(kernel):
    announce(f"test_fluxrings_sample_reproduces_the_reference_run[kernel {kernel}]")
    rows = golden_rows()
    gp = pcmx.make_params(32, ldt=True, kernel=kernel, **FLUXRINGS_KW)
    f = pcmx.farray_from_interior(pcmx_ics.fluxrings_state(32))
    pcmx.initialize_gpu(f, gp, nt=1)
    try:
        t, dt = 0.0, 0.0
        for it in range(9):
            t1, dt = pcmx.time_step(1, t, dt, ldiag=True)
            red = pcmx.get_gpu_reduced_vars()
            if it % 2 == 0:
                ours = dict(t=t, urms=red[0], umax=red[1], brms=red[2])
                announce(f"  it={it}: " + " ".join(f"{k}={float(ours[k])!r}" for k in ours))
                check_row(ours, rows[it // 2])
            t = t1
    finally:
        pcmx.finalize_gpu()
