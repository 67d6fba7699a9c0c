"""The C-ABI library loads and exports every entry point declared in
include/*.h (no compute calls: runs without a GPU)."""
import ctypes
import glob
import os
import re

from conftest import ROOT, load_package


def declared_functions():
    names = []
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        src = open(h).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        names += re.findall(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_]*\s*\*?\s*(ocfd_[a-z0-9_]+)\s*\(", src, flags=re.M)
    return sorted(set(names))


def test_header_declares_entry_points():
    names = declared_functions()
    for must in ("ocfd_init", "ocfd_flow_time_advance", "ocfd_du_inviscous", "ocfd_du_viscous", "ocfd_bc",
                 "ocfd_time_adv_rk3", "ocfd_comput_flow_variables", "ocfd_set_mesh"):
        assert must in names


def test_library_exports_every_declared_symbol():
    m = load_package()
    assert os.path.exists(m.LIB_PATH), "build the HIP library first (__graft_entry__.build())"
    lib = ctypes.CDLL(m.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_cfg_struct_layout_matches_header():
    """The ctypes mirror of ocfd_cfg has the size the C compiler gives it."""
    import subprocess
    import tempfile
    m = load_package()
    src = '#include <stdio.h>\n#include "ocfd_hip.h"\nint main(){printf("%zu", sizeof(ocfd_cfg));return 0;}\n'
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "s.c")
        open(c, "w").write(src)
        exe = os.path.join(d, "s")
        subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), c, "-o", exe], check=True)
        size = int(subprocess.run([exe], capture_output=True, text=True, check=True).stdout)
    assert size == ctypes.sizeof(m.OcfdCfg)
