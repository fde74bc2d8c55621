"""Snapshots written from the device (pcmx_write_snapshot) against the
f-array that copy_farray_from_GPU returns for the same state, in the
reference's var.dat layout (src/io_dist.f90:108-199): record 1 the evolved
variables with periodic ghosts, bitwise equal to the copied f-array; record
2 t, the periodic equidistant grid of src/grid.f90:110-216 and dx, dy, dz."""
import math

import numpy as np
import pytest

import pcmx
import pcmx_ics
import pcmx_io as I

pytestmark = pytest.mark.gpu


def announce(msg):
    print(msg, flush=True)


def grid_reference(x0, L, n):
    """x(i) as construct_grid forms it for a linear periodic grid:
    x00 + Lx*(xi - xi_lo)/(xi_up - xi_lo), xi = i - nghost - 1 + 1/2."""
    xi = np.arange(1, n + 7) - 3 - 1 + 0.5
    return x0 + L * (xi - 0.0) / (n - 0.0)


def snapshot_vs_copy(tmp_path, n, lmagnetic, steps, kernel=0, ldt=False, substeps_only=None):
    gp = pcmx.make_params(n, lmagnetic=lmagnetic, nu=5e-3, eta=5e-3, kernel=kernel, ldt=ldt, cdt=0.4)
    f0 = pcmx_ics.config_state("shells" if lmagnetic else "noise", n, 4, lmagnetic=lmagnetic)
    f = pcmx.farray_from_interior(f0)
    del f0
    pcmx.initialize_gpu(f, gp, nt=1)
    try:
        t, dt = 0.0, 0.0 if ldt else 1e-3
        if substeps_only:
            for s in range(1, substeps_only + 1):
                dt = pcmx.rhs_gpu(s, t, dt)
        elif steps:
            t, dt = pcmx.time_step(steps, t, dt)
        path = tmp_path / "var.dat"
        xyz0 = (-math.pi, -1.0, 0.25)
        pcmx.write_snapshot(str(path), t, xyz0)
        copy = np.zeros_like(f)
        pcmx.copy_farray_from_gpu(copy)
    finally:
        pcmx.finalize_gpu()
    return gp, path, copy, t, xyz0


@pytest.mark.parametrize("kernel", [1, 3])
def test_snapshot_equals_copied_farray(tmp_path, kernel):
    n = (32, 16, 24)
    announce(f"test_snapshot_equals_copied_farray[kernel {kernel}] n={n}")
    gp, path, copy, t, xyz0 = snapshot_vs_copy(tmp_path, n, True, 3, kernel=kernel)
    g, tt, x, y, z, dx, dy, dz = I.read_var(str(path), gp.mvar, n)
    np.testing.assert_array_equal(g, copy[:gp.mvar])
    assert tt == t and (dx, dy, dz) == (gp.dx, gp.dy, gp.dz)
    for c, x0, m, d in ((x, xyz0[0], n[0], gp.dx), (y, xyz0[1], n[1], gp.dy), (z, xyz0[2], n[2], gp.dz)):
        np.testing.assert_array_equal(c, x0 + (np.arange(1, m + 7) - 3.5) * d)
        ref = grid_reference(x0, m * d, m)
        assert np.abs(c - ref).max() <= 4 * np.finfo(float).eps * max(1.0, np.abs(ref).max())
    # the periodic ghost zones of the file repeat the interior
    np.testing.assert_array_equal(g[:, 3:-3, 3:-3, 0:3], g[:, 3:-3, 3:-3, n[0]:n[0] + 3])


def test_snapshot_between_deferred_substeps(tmp_path):
    """Courant steps of the split kernel keep f1 implicit after substep 1; a
    snapshot there holds f1, as the copied f-array does."""
    n = (32, 16, 24)
    announce("test_snapshot_between_deferred_substeps")
    gp, path, copy, t, _ = snapshot_vs_copy(tmp_path, n, True, 0, kernel=3, ldt=True, substeps_only=1)
    g = I.read_var(str(path), gp.mvar, n)[0]
    np.testing.assert_array_equal(g, copy[:gp.mvar])


def test_snapshot_record_over_2gib(tmp_path):
    """A 352^3 hydro state (358^3 x 4 x 8 bytes = 1.47 GB) is below the
    subrecord limit; 7-variable MHD at 352^3 (2.57 GB) is split into two
    subrecords."""
    n = 352
    announce(f"test_snapshot_record_over_2gib n={n}")
    gp, path, copy, t, _ = snapshot_vs_copy(tmp_path, n, True, 0)
    with open(path, "rb") as fh:
        lead = np.frombuffer(fh.read(4), "<i4")[0]
    assert lead == -2147483639
    g = I.read_var(str(path), gp.mvar, n)[0]
    assert np.array_equal(g, copy[:gp.mvar])
