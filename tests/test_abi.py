"""CPU-side checks of the drop-in boundary: the C-ABI library loads, exports
every entry point include/pcmx.h declares, and rejects bad calls with the
documented error codes before touching a device."""
import ctypes
import os
import re

import pytest

import pcmx

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "pcmx.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(pcmx_\w+)\s*\(", text, re.M)))


def test_header_declares_the_module_gpu_entry_points():
    syms = declared_symbols()
    for s in ("pcmx_initialize", "pcmx_finalize", "pcmx_load_farray", "pcmx_copy_farray", "pcmx_rhs",
              "pcmx_before_boundary", "pcmx_update_after_substep", "pcmx_prepare_for_first_substep",
              "pcmx_get_reduced_vars", "pcmx_update_by_name", "pcmx_test_bcs", "pcmx_split_update"):
        assert s in syms


def test_library_exports_every_declared_symbol():
    if not os.path.exists(pcmx.LIB_PATH):
        pytest.fail("libpcmx.so not built: run __graft_entry__.build()")
    lib = ctypes.CDLL(pcmx.LIB_PATH)
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_calls_before_initialize_fail_with_state_error():
    lib = pcmx.lib()
    d = ctypes.c_double(1e-3)
    assert lib.pcmx_rhs(1, 0.0, ctypes.byref(d)) == -2
    assert b"pcmx_initialize" in lib.pcmx_last_error()
    assert lib.pcmx_copy_farray(None, 4) == -2
    with pytest.raises(pcmx.PcmxError):
        pcmx.get_gpu_reduced_vars()


@pytest.mark.parametrize("mutate,msg", [
    (lambda p: setattr(p, "mvar", 5), b"mvar"),
    (lambda p: setattr(p, "iax", 6), b"slots"),
    (lambda p: setattr(p, "nx", 2), b"3 interior"),
    (lambda p: setattr(p, "dx", 0.0), b"spacing"),
    (lambda p: setattr(p, "precision", 3), b"precision"),
    (lambda p: setattr(p, "kernel", 9), b"kernel"),
])
def test_initialize_rejects_bad_parameters(mutate, msg):
    p = pcmx.make_params(16, lmagnetic=True)
    mutate(p)
    lib = pcmx.lib()
    assert lib.pcmx_initialize(None, 0, ctypes.byref(p), 0.0, 0) == -1
    assert msg in lib.pcmx_last_error()


def test_entropy_requires_gamma_not_one():
    p = pcmx.make_params(16, lentropy=True, gamma=1.0)
    assert pcmx.lib().pcmx_initialize(None, 0, ctypes.byref(p), 0.0, 0) == -1


def test_ensemble_calls_before_create_fail_with_state_error():
    lib = pcmx.lib()
    assert lib.pcmx_ens_step(1, 0) == -2
    assert b"pcmx_ens_create" in lib.pcmx_last_error()


@pytest.mark.parametrize("bad", ["untiled", "entropy", "courant", "no_dt"])
def test_ensemble_rejects_unsupported_setups_before_allocating(bad):
    par = pcmx.make_params(32, lmagnetic=True, lentropy=(bad == "entropy"), gamma=5 / 3 if bad == "entropy" else 1.0,
                           ldt=(bad == "courant"), eta=1e-3)
    boxes = [pcmx.make_box(32, 1e-3, 1e-3, 1e-3), pcmx.make_box(24 if bad == "untiled" else 48, 1e-3, 1e-3,
                                                                0.0 if bad == "no_dt" else 1e-3)]
    with pytest.raises(pcmx.PcmxError):
        pcmx.ens_create(par, boxes)
