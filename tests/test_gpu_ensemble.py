"""Ensembles of independent boxes (pcmx_ens_*): every box must evolve exactly
as it does alone in the single-box engine, whatever tile shape, chunk length
and launch mode (HIP graph or plain launches) the ensemble uses.  Tolerances:
FP64 relative L-inf per variable <= 1e-12 against the direct single-box
kernel and the oracle; the mixed form <= 1e-5 against FP64 after 20 steps."""
import math

import numpy as np
import pytest

import oracle as O
import pcmx
import pcmx_ics

pytestmark = pytest.mark.gpu


def rel_linf(a, b):
    out = []
    for v in range(a.shape[0]):
        den = np.abs(b[v]).max()
        num = np.abs(a[v] - b[v]).max()
        out.append(num / den if den > 0 else num)
    return np.array(out)


def make_ensemble(sizes, mag=True, seed=5):
    rng = np.random.default_rng(seed)
    boxes, states = [], []
    for i, n in enumerate(sizes):
        eta = float(10 ** rng.uniform(-3, -2))
        nu = float(eta * 10 ** rng.uniform(math.log10(0.5), 1.0))
        dt = 1e-3 if n >= 64 else 2e-3
        boxes.append(pcmx.make_box(n, nu, eta if mag else 0.0, dt))
        states.append(pcmx_ics.config_state("shells", n, seed + i, lmagnetic=mag))
    return boxes, states


def single_box(state, box, mag, steps, kernel=1):
    n = (box.nx, box.ny, box.nz)
    gp = pcmx.make_params(n, lmagnetic=mag, nu=box.nu, eta=box.eta, kernel=kernel,
                          L=(box.dx * box.nx, box.dy * box.ny, box.dz * box.nz))
    g, _, _ = pcmx.advance(state, gp, steps, box.dt)
    return g


def run_ensemble(boxes, states, mag, steps, flags=0, precision=pcmx.PREC_FP64, tune=False, ldiag=True):
    par = pcmx.make_params(32, lmagnetic=mag, precision=precision)
    pcmx.ens_create(par, boxes, flags)
    try:
        for i, f in enumerate(states):
            pcmx.ens_load(i, f)
        if tune:
            pcmx.ens_tune(1)
        pcmx.ens_step(steps, ldiag=ldiag)
        out = [pcmx.ens_copy(i, f.shape) for i, f in enumerate(states)]
        red = pcmx.ens_reduced(len(boxes))
        info = pcmx.ens_timing()
    finally:
        pcmx.ens_destroy()
    return out, red, info


SIZES = [32, 48, 32, 64, 48]


@pytest.mark.parametrize("mag", [True, False])
def test_ensemble_matches_single_boxes(mag):
    boxes, states = make_ensemble(SIZES, mag=mag)
    print(f"test_ensemble_matches_single_boxes: mag={mag} sizes={SIZES}", flush=True)
    out, _, info = run_ensemble(boxes, states, mag, 3)
    print(f"  plan (ms, substeps, tile, lz, bpc, items) = {info}", flush=True)
    for i, (b, f) in enumerate(zip(boxes, states)):
        ref = single_box(f, b, mag, 3)
        err = rel_linf(out[i], ref)
        print(f"  box {i} n={b.nx}: {err.max():.2e}", flush=True)
        assert np.all(err <= 1e-12), (i, err)


def test_ensemble_one_step_matches_oracle():
    boxes, states = make_ensemble([48, 32], mag=True, seed=9)
    print("test_ensemble_one_step_matches_oracle", flush=True)
    out, red, _ = run_ensemble(boxes, states, True, 1)
    for i, (b, f) in enumerate(zip(boxes, states)):
        op = O.make_params(b.nx, lmagnetic=True, nu=b.nu, eta=b.eta, dt=b.dt)
        o, _, _, _ = O.advance(f, op, 1)
        assert np.all(rel_linf(out[i], o) <= 1e-12)
        _, _, (urms, umax, brms) = O.rhs(f, op)
        np.testing.assert_allclose(red[i], [urms, umax, brms], rtol=1e-13)


@pytest.mark.parametrize("steps", [4, 5])
def test_graph_and_plain_launches_agree(steps):
    boxes, states = make_ensemble([32, 48, 64], mag=True, seed=3)
    print(f"test_graph_and_plain_launches_agree: {steps} steps", flush=True)
    a, ra, _ = run_ensemble(boxes, states, True, steps)
    b, rb, _ = run_ensemble(boxes, states, True, steps, flags=pcmx.ENS_NO_GRAPH)
    for x, y in zip(a, b):
        np.testing.assert_array_equal(x, y)
    np.testing.assert_array_equal(ra, rb)


def test_tuned_plan_gives_the_same_result():
    boxes, states = make_ensemble([32, 48, 96], mag=True, seed=4)
    print("test_tuned_plan_gives_the_same_result", flush=True)
    a, _, ia = run_ensemble(boxes, states, True, 2)
    b, _, ib = run_ensemble(boxes, states, True, 2, tune=True)
    print(f"  default plan {ia[2:]}, tuned plan {ib[2:]}", flush=True)
    for x, y in zip(a, b):
        assert np.all(rel_linf(y, x) <= 1e-13)


def test_more_work_items_than_blocks():
    sizes = [64] * 12
    boxes, states = make_ensemble(sizes, mag=True, seed=21)
    print(f"test_more_work_items_than_blocks: {len(sizes)} boxes of 64^3", flush=True)
    out, _, info = run_ensemble(boxes, states, True, 2)
    print(f"  plan {info}", flush=True)
    assert info[5] > 256
    for i in (0, 5, 11):
        ref = single_box(states[i], boxes[i], True, 2)
        assert np.all(rel_linf(out[i], ref) <= 1e-12)


def test_mixed_ensemble_error_after_20_steps():
    boxes, states = make_ensemble([32, 48, 64], mag=True, seed=8)
    print("test_mixed_ensemble_error_after_20_steps", flush=True)
    a, _, _ = run_ensemble(boxes, states, True, 20)
    b, _, _ = run_ensemble(boxes, states, True, 20, precision=pcmx.PREC_MIXED)
    for x, y in zip(a, b):
        err = rel_linf(y, x)
        print(f"  {err.max():.2e}", flush=True)
        assert np.all(err <= 1e-5), err
