"""The Fortran side of the drop-in boundary compiles with flang:
pcmx_iface.f90 (iso_c_binding interfaces of include/pcmx.h) always, and the
module GPU replacement gpu_mi355x.f90 against the reference's own
src/gpu.h public list (read in place) plus interface stubs of the Cdata /
Messages / Sub ... names it uses, when the reference tree is present."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FC = "/opt/rocm/lib/llvm/bin/flang"
FDIR = os.path.join(ROOT, "pencil-code_amd", "fortran")
REF_SRC = "/root/reference/src"


def flang(args, cwd):
    return subprocess.run([FC, "-fdefault-real-8", "-fdefault-double-8"] + args, cwd=cwd,
                          capture_output=True, text=True, timeout=300)


@pytest.mark.skipif(not os.path.exists(FC), reason="flang not installed")
def test_iface_and_driver_compile(tmp_path):
    r = flang(["-c", os.path.join(FDIR, "pcmx_iface.f90"), os.path.join(FDIR, "pcmx_run.f90")], str(tmp_path))
    assert r.returncode == 0, r.stderr


@pytest.mark.skipif(not (os.path.exists(FC) and os.path.exists(os.path.join(REF_SRC, "gpu.h"))),
                    reason="needs flang and the reference's src/gpu.h")
def test_module_gpu_compiles_against_gpu_h(tmp_path):
    r = flang(["-c", os.path.join(ROOT, "tests", "fortran_stubs", "stubs.f90")], str(tmp_path))
    assert r.returncode == 0, r.stderr
    r = flang(["-c", os.path.join(FDIR, "pcmx_iface.f90")], str(tmp_path))
    assert r.returncode == 0, r.stderr
    r = flang(["-I" + REF_SRC, "-c", os.path.join(FDIR, "gpu_mi355x.f90")], str(tmp_path))
    assert r.returncode == 0, r.stderr
    assert os.path.exists(os.path.join(str(tmp_path), "gpu.mod"))
