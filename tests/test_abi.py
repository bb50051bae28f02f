"""The C ABI: the library loads without a GPU, exports every entry point the
header declares, and the ctypes mirrors have the C layout."""
import ctypes as C
import re
import subprocess
from pathlib import Path

import pytest

from pytnl_lbm import _abi

ROOT = Path(__file__).resolve().parent.parent
HEADER = ROOT / "include" / "tnl_lbm_amd.h"


def declared_functions():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(tnl_lbm_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = _abi.load_library()
    names = declared_functions()
    assert "tnl_lbm_d3q27_step" in names and len(names) >= 10
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert lib.tnl_lbm_abi_version() == _abi.ABI_VERSION


def test_struct_layout_matches_c(tmp_path):
    src = tmp_path / "sz.c"
    src.write_text(f'#include "{HEADER}"\n#include <stdio.h>\n#include <stddef.h>\n'
                   "int main(void){printf(\"%zu %zu %zu %zu %zu\\n\", sizeof(tnl_lbm_config), "
                   "sizeof(tnl_lbm_d3q27_data), sizeof(tnl_lbm_d3q7_data), "
                   "offsetof(tnl_lbm_d3q27_data, inflow_profile), offsetof(tnl_lbm_d3q7_data, inflow_phi));return 0;}\n")
    exe = tmp_path / "sz"
    subprocess.run(["gcc", "-o", str(exe), str(src)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()
    got = [C.sizeof(_abi.Config), C.sizeof(_abi.D3Q27Data), C.sizeof(_abi.D3Q7Data),
           _abi.D3Q27Data.inflow_profile.offset, _abi.D3Q7Data.inflow_phi.offset]
    assert [int(v) for v in out] == got


def test_validation_errors_without_gpu():
    """Argument validation runs before any HIP call, so it is testable on the CPU."""
    lib = _abi.load_library()
    cfg = _abi.Config.make()
    d = _abi.D3Q27Data.make(4, 4, 4)
    off, end = _abi.I64x3(0, 0, 0), _abi.I64x3(4, 4, 4)
    assert lib.tnl_lbm_d3q27_step(None, C.byref(d), off, end, 0, None) == _abi.ERR_NULL
    assert lib.tnl_lbm_d3q27_step(C.byref(cfg), C.byref(d), off, end, 0, None) == _abi.ERR_NULL  # no arrays
    d.df[0] = 0x1000
    bad = _abi.Config.make()
    bad.struct_size = 4
    assert lib.tnl_lbm_d3q27_step(C.byref(bad), C.byref(d), off, end, 0, None) == _abi.ERR_VERSION
    assert b"struct_size" in lib.tnl_lbm_last_error()
    d.XYZ = 10
    assert lib.tnl_lbm_d3q27_step(C.byref(cfg), C.byref(d), off, end, 0, None) == _abi.ERR_DIMS
    d.XYZ = 64
    assert lib.tnl_lbm_d3q27_step(C.byref(cfg), C.byref(d), off, _abi.I64x3(5, 4, 4), 0, None) == _abi.ERR_DIMS
    d.lbm_viscosity = 0.0
    assert lib.tnl_lbm_d3q27_step(C.byref(cfg), C.byref(d), off, end, 0, None) == _abi.ERR_ARG
    cfg2 = _abi.Config.make(collision=_abi.COLL_CUM_2017, storage=_abi.STORE_F16_DEV)
    d.lbm_viscosity = 0.01
    assert lib.tnl_lbm_d3q27_step(C.byref(cfg2), C.byref(d), off, end, 0, None) == _abi.ERR_UNSUPPORTED


def test_halo_validation_without_gpu():
    """tnl_lbm_d3q27_halo_x rejects slabs without ghost planes, mismatched
    faces, unknown phases and the push phase of A-B before any copy."""
    lib = _abi.load_library()
    cfg = _abi.Config.make()
    lo = _abi.D3Q27Data.make(4, 6, 5, 1, 0, 0)
    up = _abi.D3Q27Data.make(3, 6, 5, 1, 0, 0)
    lo.df[0] = up.df[0] = 0x1000
    flat = _abi.D3Q27Data.make(4, 6, 5)
    flat.df[0] = 0x1000
    assert lib.tnl_lbm_d3q27_halo_x(C.byref(cfg), C.byref(lo), C.byref(flat), _abi.HALO_PULL, None) == _abi.ERR_DIMS
    other = _abi.D3Q27Data.make(4, 7, 5, 1, 0, 0)
    other.df[0] = 0x1000
    assert lib.tnl_lbm_d3q27_halo_x(C.byref(cfg), C.byref(lo), C.byref(other), _abi.HALO_PULL, None) == _abi.ERR_DIMS
    assert lib.tnl_lbm_d3q27_halo_x(C.byref(cfg), C.byref(lo), C.byref(up), 7, None) == _abi.ERR_ARG
    ab = _abi.Config.make(pattern=_abi.PATTERN_AB)
    lo.df[1] = up.df[1] = 0x2000
    assert lib.tnl_lbm_d3q27_halo_x(C.byref(ab), C.byref(lo), C.byref(up), _abi.HALO_PUSH, None) == _abi.ERR_UNSUPPORTED
    assert lib.tnl_lbm_d3q27_halo_x(None, C.byref(lo), C.byref(up), _abi.HALO_PULL, None) == _abi.ERR_NULL


def test_adjoint_rejects_a_force_field_without_gpu():
    """The adjoint entry points differentiate the homogeneous-force step only:
    a per-site force field is refused before any launch."""
    lib = _abi.load_library()
    cfg = _abi.Config.make(pattern=_abi.PATTERN_AB)
    d = _abi.D3Q27Data.make(4, 4, 4)
    d.df[0] = d.df[1] = 0x1000
    d.lbm_viscosity = 0.01
    d.force_field = 0x3000
    assert lib.tnl_lbm_d3q27_adjoint_step(C.byref(cfg), C.byref(d), C.c_void_p(0x4000), C.c_void_p(0x5000), None,
                                          None) == _abi.ERR_UNSUPPORTED
    assert b"force_field" in lib.tnl_lbm_last_error()
    assert lib.tnl_lbm_d3q27_step_record(C.byref(cfg), C.byref(d), C.c_void_p(0x6000), 0, None) == _abi.ERR_UNSUPPORTED
