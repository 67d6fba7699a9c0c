"""The Fortran side of the drop-in boundary (CPU only, no GPU calls).

Compiles opencfd-comb_amd/fortran/ocfd_hip.f90 (the iso_c_binding module the
reference's Fortran driver uses) with flang, checks that the bind(C) type
ocfd_cfg has the C struct's size and field offsets, and links and runs a
Fortran caller against libocfd_hip.so: ocfd_init with LAP = 3 must return a
non-zero status and ocfd_last_error must come back through the shim as the
library's message (the reference's print + stop path).  The error is raised
by argument checking, before any HIP call, so no GPU is needed.
"""
import os
import shutil
import subprocess

import pytest

from conftest import ROOT

FLANG = shutil.which("flang") or "/opt/rocm/lib/llvm/bin/flang"
SHIM = os.path.join(ROOT, "opencfd-comb_amd", "fortran", "ocfd_hip.f90")
LIBDIR = os.path.join(ROOT, "opencfd-comb_amd")

FIELDS = ["nx", "ny", "nz", "lap", "n_spec", "periodic", "scheme_invis", "scheme_vis", "flux_splitting",
          "if_viscous", "if_scheme_character", "scheme_boundary", "ghost_cell", "ibc", "if_mass_force",
          "if_real_gas", "flag_vis", "flag_cond", "flag_diff", "wall_catalysis", "pad0", "ud7l_diss",
          "hybrid_para", "amu0", "pr", "sc", "mass_force", "bc_para"]

F_PROG = """
program shim_check
  use iso_c_binding
  use ocfd_hip
  implicit none
  type(ocfd_cfg), target :: c
  integer(c_int32_t) :: rc
  print '(A,I0)', 'SIZE ', c_sizeof(c)
{offsets}
  c%nx = 16; c%ny = 16; c%nz = 16; c%lap = 3; c%n_spec = 1
  rc = ocfd_init(c)
  print '(A,I0)', 'RC ', rc
  print '(A,A)', 'ERR ', ocfd_error_text()
end program
"""

C_PROG = """
#include <stddef.h>
#include <stdio.h>
#include "ocfd_hip.h"
int main(void) {{
  printf("SIZE %zu\\n", sizeof(ocfd_cfg));
{offsets}
  return 0;
}}
"""


def _run(cmd, cwd):
    r = subprocess.run(cmd, cwd=cwd, capture_output=True, text=True)
    assert r.returncode == 0, (cmd, r.stdout, r.stderr)
    return r.stdout


@pytest.fixture(scope="module")
def built(tmp_path_factory):
    if not os.path.exists(FLANG):
        pytest.skip("no Fortran compiler")
    if not os.path.exists(os.path.join(LIBDIR, "libocfd_hip.so")):
        pytest.skip("library not built")
    d = tmp_path_factory.mktemp("fshim")
    f_off = "\n".join(f"  print '(A,I0)', 'OFF {n} ', loc(c%{n}) - loc(c)" for n in FIELDS)
    (d / "check.f90").write_text(F_PROG.format(offsets=f_off))
    c_off = "\n".join(f'  printf("OFF {n} %zu\\n", offsetof(ocfd_cfg, {n}));' for n in FIELDS)
    (d / "check.c").write_text(C_PROG.format(offsets=c_off))
    _run([FLANG, "-c", SHIM, "-o", "ocfd_hip.o"], d)
    _run([FLANG, "check.f90", "ocfd_hip.o", "-o", "fcheck", f"-L{LIBDIR}", "-locfd_hip",
          f"-Wl,-rpath,{LIBDIR}", "-L/opt/rocm/lib", "-Wl,-rpath,/opt/rocm/lib"], d)
    _run(["gcc", "-I" + os.path.join(ROOT, "include"), "check.c", "-o", "ccheck"], d)
    return d


def _parse(text):
    out = {}
    for line in text.splitlines():
        parts = line.split()
        if not parts:
            continue
        if parts[0] == "OFF":
            out["OFF " + parts[1]] = int(parts[2])
        elif parts[0] in ("SIZE", "RC"):
            out[parts[0]] = int(parts[1])
        elif parts[0] == "ERR":
            out["ERR"] = line[4:].strip()
    return out


def test_cfg_layout_matches_c(built):
    f = _parse(_run(["./fcheck"], built))
    c = _parse(_run(["./ccheck"], built))
    assert f["SIZE"] == c["SIZE"] == 576
    for n in FIELDS:
        assert f["OFF " + n] == c["OFF " + n], n


def test_fortran_caller_gets_status_and_message(built):
    f = _parse(_run(["./fcheck"], built))
    assert f["RC"] != 0
    assert "LAP must be 4" in f["ERR"], f["ERR"]
