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


LAYOUT_F90 = """
program layout
  use iso_c_binding
  use pcmx_iface
  implicit none
  type(pcmx_params) :: p
  type(pcmx_forcing) :: fr
  integer(c_int8_t), allocatable :: b(:)
  integer :: u, k
  p%nx = 11; p%ny = 12; p%nz = 13; p%mvar = 7
  p%iux = 0; p%ilnrho = 3; p%iss = -1; p%iax = 4
  p%dx = 0.5d0; p%dy = 0.25d0; p%dz = 0.125d0
  p%nu = 1.5d-3; p%eta = 2.5d-3; p%chi = 3.5d-3
  p%cs0 = 1.1d0; p%gamma = 1.2d0; p%cp = 1.3d0; p%rho0 = 1.4d0; p%mu0 = 1.6d0
  p%cdt = 0.7d0; p%cdtv = 0.3d0
  p%ldt = 1; p%precision = 1; p%kernel = 3
  p%lupw_uu = 1; p%lupw_lnrho = 0
  p%gravz = -0.75d0; p%lperi_z = 0; p%lupw_ss = 1
  do k = 1, 8
    p%bcz_bot(k) = k; p%bcz_top(k) = 10 + k
  enddo
  fr%active = 1; fr%target = 4; fr%scale_sqrt_dt = 1; fr%reserved = 0
  fr%kk = [1d0, 2d0, 3d0]; fr%phase = 0.25d0; fr%fact = 0.5d0
  fr%coef1 = [4d0, 5d0, 6d0]; fr%coef2 = [7d0, 8d0, 9d0]; fr%fda = [1d0, 1.5d0, 2d0]
  fr%k1_ff = 2d0; fr%xyz0 = [-1d0, -2d0, -3d0]
  open(newunit=u, file='layout.bin', access='stream', form='unformatted', status='replace')
  b = transfer(p, [0_c_int8_t])
  write(u) size(b, kind=c_int64_t)
  write(u) b
  b = transfer(fr, [0_c_int8_t])
  write(u) size(b, kind=c_int64_t)
  write(u) b
  close(u)
end program layout
"""

LAYOUT_C = r"""
#include <stddef.h>
#include <stdio.h>
#include "pcmx.h"
int main(void) {
  printf("%zu %zu %zu %zu %zu %zu\n", sizeof(pcmx_params), offsetof(pcmx_params, gravz),
         offsetof(pcmx_params, lperi_z), offsetof(pcmx_params, bcz_top), sizeof(pcmx_forcing),
         offsetof(pcmx_forcing, xyz0));
  return 0;
}
"""


@pytest.mark.skipif(not os.path.exists(FC), reason="flang not installed")
def test_bind_c_layouts_match_the_c_header(tmp_path):
    """The Fortran bind(C) types, the C header and the ctypes mirrors agree
    byte for byte (sizes, padding, field offsets)."""
    import ctypes
    import sys
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "pencil-code_amd"))
    import pcmx
    src = tmp_path / "layout.f90"
    src.write_text(LAYOUT_F90)
    lib = os.path.join(ROOT, "pencil-code_amd", "lib")
    if not os.path.exists(os.path.join(lib, "libpcmx.so")):
        pytest.skip("libpcmx.so not built")
    r = flang(["-o", "layout", os.path.join(FDIR, "pcmx_iface.f90"), str(src), f"-L{lib}", "-lpcmx",
               f"-Wl,-rpath,{lib}", "-Wl,-rpath,/opt/rocm/lib"], str(tmp_path))
    assert r.returncode == 0, r.stderr
    r = subprocess.run(["./layout"], cwd=str(tmp_path), capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    raw = (tmp_path / "layout.bin").read_bytes()
    n1 = int(np.frombuffer(raw[:8], "<i8")[0])
    pb, rest = raw[8:8 + n1], raw[8 + n1:]
    n2 = int(np.frombuffer(rest[:8], "<i8")[0])
    fb = rest[8:8 + n2]
    assert n1 == ctypes.sizeof(pcmx.Params) and n2 == ctypes.sizeof(pcmx.Forcing)
    p = pcmx.Params.from_buffer_copy(pb)
    assert (p.nx, p.ny, p.nz, p.mvar, p.iax, p.kernel, p.lupw_uu, p.lperi_z) == (11, 12, 13, 7, 4, 3, 1, 0)
    assert (p.dz, p.mu0, p.cdtv, p.gravz, p.lupw_ss) == (0.125, 1.6, 0.3, -0.75, 1)
    assert list(p.bcz_bot) == list(range(1, 9)) and list(p.bcz_top) == list(range(11, 19))
    fr = pcmx.Forcing.from_buffer_copy(fb)
    assert (fr.active, fr.target, fr.scale_sqrt_dt) == (1, 4, 1)
    assert list(fr.kk) == [1, 2, 3] and list(fr.fda) == [1, 1.5, 2] and list(fr.xyz0) == [-1, -2, -3]
    assert (fr.phase, fr.fact, fr.k1_ff) == (0.25, 0.5, 2.0)
    # the C compiler's view of the header
    csrc = tmp_path / "layout.c"
    csrc.write_text(LAYOUT_C)
    r = subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), "-o", "layout_c", str(csrc)], cwd=str(tmp_path),
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    out = subprocess.run(["./layout_c"], cwd=str(tmp_path), capture_output=True, text=True, timeout=60).stdout.split()
    assert [int(v) for v in out] == [ctypes.sizeof(pcmx.Params), pcmx.Params.gravz.offset, pcmx.Params.lperi_z.offset,
                                     pcmx.Params.bcz_top.offset, ctypes.sizeof(pcmx.Forcing), pcmx.Forcing.xyz0.offset]
