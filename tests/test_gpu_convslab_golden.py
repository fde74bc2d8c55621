"""The samples/conv-slab run of the reference (reference.out.double, 32^3,
40 Courant steps) on the HIP path through the C ABI, on the split kernel
(the automatic choice for this sample) and on the direct kernel: t, dt,
urms, umax, rhom and ssm at it = 0, 10, 20, 30, 40 against the reference's
printed values (about 12 significant digits for urms and umax), with the
same tolerances as the oracle's golden test."""
import numpy as np
import pytest

import oracle as O
import pcmx
import pcmx_ics
from test_gpu_parity import announce
from test_oracle_convslab_golden import (CONV_SLAB_KW, VARIANTS, check_row, conv_slab_layers, golden_rows,
                                         means_after_bcs)

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("kernel,variant", [(0, "cp1"), (1, "cp1"), (0, "cp2")])
def test_conv_slab_sample_reproduces_the_reference_run(kernel, variant):
    announce(f"test_conv_slab_sample_reproduces_the_reference_run[kernel {kernel}, {variant}]")
    rows = golden_rows(variant)
    cp = VARIANTS[variant][1]
    gp = pcmx.make_params((32, 32, 32), ldt=True, kernel=kernel, cp=cp, layers=conv_slab_layers(pcmx, variant),
                          **CONV_SLAB_KW)
    op = O.make_params((32, 32, 32), dt=0.0, cp=cp, layers=conv_slab_layers(O, variant), **CONV_SLAB_KW)
    f0 = pcmx_ics.conv_slab_state(32, cp=cp)
    f = pcmx.farray_from_interior(f0)
    pcmx.initialize_gpu(f, gp, nt=1)
    try:
        t, dt = 0.0, 0.0
        state = f0
        for it in range(41):
            t1, dt = pcmx.time_step(1, t, dt, ldiag=True)
            red = pcmx.get_gpu_reduced_vars()
            if it % 10 == 0:
                rhom, ssm = means_after_bcs(state, op)
                ours = dict(t=t, dt=dt, urms=red[0], umax=red[1], rhom=rhom, ssm=ssm)
                announce(f"  it={it}: " + " ".join(f"{k}={ours[k]!r}" for k in ours))
                check_row(ours, rows[it // 10])
            if it % 10 == 9:
                out = np.zeros_like(f)
                pcmx.copy_farray_from_gpu(out)
                state = pcmx.interior_of(out, 5).copy()
            t = t1
    finally:
        pcmx.finalize_gpu()
