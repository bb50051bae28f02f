"""ctypes mirror of include/tnl_lbm_amd.h and the loader of libtnl_lbm_amd.so.

The structs here must match the C header field by field; tests/test_abi.py
checks the sizes against the compiled library's expectations (every entry
point rejects a struct whose ``struct_size`` differs).
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

ABI_VERSION = 4

# status codes
OK = 0
ERR_NULL = -1
ERR_DIMS = -2
ERR_UNSUPPORTED = -3
ERR_VERSION = -4
ERR_HIP = -5
ERR_ARG = -6

# enum tnl_lbm_storage
STORE_F64 = 0
STORE_F32 = 1
STORE_F32_DEV = 2
STORE_F16_DEV = 3
STORE_BF16_DEV = 4
STORE_I16_DEV = 5
STORAGE_NAMES = {"f64": STORE_F64, "f32": STORE_F32, "f32dev": STORE_F32_DEV,
                 "f16dev": STORE_F16_DEV, "bf16dev": STORE_BF16_DEV, "i16dev": STORE_I16_DEV}
STORAGE_BYTES = {STORE_F64: 8, STORE_F32: 4, STORE_F32_DEV: 4, STORE_F16_DEV: 2, STORE_BF16_DEV: 2,
                 STORE_I16_DEV: 2}

# enum tnl_lbm_real
REAL_F64 = 0
REAL_F32 = 1

# enum tnl_lbm_pattern
PATTERN_AA = 0
PATTERN_AB = 1

# enum tnl_lbm_collision
COLL_CUM = 0
COLL_CUM_2017 = 1
COLL_CUM_2017_ANTIALIAS = 2
COLL_MRT_LES = 3
COLL_KBC_N1, COLL_KBC_N2, COLL_KBC_N3, COLL_KBC_N4 = 4, 5, 6, 7
COLL_KBC_C1, COLL_KBC_C2, COLL_KBC_C3, COLL_KBC_C4 = 8, 9, 10, 11
COLLISION_NAMES = {"CUM": COLL_CUM, "CUM_2017": COLL_CUM_2017,
                   "CUM_2017_ANTIALIAS": COLL_CUM_2017_ANTIALIAS, "MRT_LES": COLL_MRT_LES,
                   "KBC_N1": COLL_KBC_N1, "KBC_N2": COLL_KBC_N2, "KBC_N3": COLL_KBC_N3,
                   "KBC_N4": COLL_KBC_N4, "KBC_C1": COLL_KBC_C1, "KBC_C2": COLL_KBC_C2,
                   "KBC_C3": COLL_KBC_C3, "KBC_C4": COLL_KBC_C4}

# enum tnl_lbm_ade_collision
ADE_COLL_SRT = 0
ADE_COLL_MRT = 1
ADE_COLL_CLBM = 2
ADE_COLLISION_NAMES = {"SRT": ADE_COLL_SRT, "MRT": ADE_COLL_MRT, "CLBM": ADE_COLL_CLBM}

FLAG_OUTPUT_MACRO = 0x1
FLAG_DISTRIBUTED_X = 0x2
FLAG_DISTRIBUTED_Y = 0x4
FLAG_DISTRIBUTED_Z = 0x8

# tnl_lbm_halo_phase
HALO_PULL = 0
HALO_PUSH = 1


class GEO:
    """D3Q27_BC_All::GEO (d3q27/bc.h:17-43)."""
    FLUID = 0
    WALL = 1
    INFLOW = 2
    INFLOW_LEFT = 3
    INFLOW_BB_LEFT = 4
    INFLOW_EQ_LEFT = 5
    OUTFLOW_EQ = 6
    OUTFLOW_RIGHT = 7
    OUTFLOW_RIGHT_INTERP = 8
    PERIODIC = 9
    NOTHING = 10
    SYM_TOP = 11
    SYM_BOTTOM = 12
    SYM_LEFT = 13
    SYM_RIGHT = 14
    SYM_BACK = 15
    SYM_FRONT = 16


class ADE_GEO:
    """D3Q7_BC_All::GEO (d3q7/bc.h:17-36)."""
    FLUID = 0
    WALL = 1
    WALL_BODY = 2
    SOLID = 3
    INFLOW = 7
    OUTFLOW_RIGHT = 8
    PERIODIC = 9
    NOTHING = 10
    SYM_TOP = 12
    SYM_BOTTOM = 13
    SYM_LEFT = 14
    SYM_RIGHT = 15
    SYM_BACK = 16
    SYM_FRONT = 17


class Config(C.Structure):
    _fields_ = [("struct_size", C.c_uint32), ("version", C.c_uint32), ("collision", C.c_int32),
                ("storage", C.c_int32), ("pattern", C.c_int32), ("macro_real", C.c_int32)]

    @classmethod
    def make(cls, collision=COLL_CUM, storage=STORE_F64, pattern=PATTERN_AA, macro_real=REAL_F64):
        return cls(C.sizeof(cls), ABI_VERSION, collision, storage, pattern, macro_real)


class D3Q27Data(C.Structure):
    _fields_ = [("struct_size", C.c_uint32), ("version", C.c_uint32),
                ("X", C.c_int64), ("Y", C.c_int64), ("Z", C.c_int64),
                ("ox", C.c_int64), ("oy", C.c_int64), ("oz", C.c_int64), ("XYZ", C.c_int64),
                ("even_iter", C.c_int32), ("inflow_profile_real", C.c_int32),
                ("lbm_viscosity", C.c_double), ("fx", C.c_double), ("fy", C.c_double), ("fz", C.c_double),
                ("df", C.c_void_p * 2), ("macro", C.c_void_p), ("map", C.c_void_p),
                ("inflow_vx", C.c_double), ("inflow_vy", C.c_double), ("inflow_vz", C.c_double),
                ("inflow_profile", C.c_void_p * 3), ("store_range", C.c_double),
                ("store_range_class", C.c_double * 4), ("force_field", C.c_void_p),
                ("rows", C.c_void_p), ("row_count", C.c_int64 * 3)]

    @classmethod
    def make(cls, X, Y, Z, ox=0, oy=0, oz=0, XYZ=None):
        d = cls()
        d.struct_size = C.sizeof(cls)
        d.version = ABI_VERSION
        d.X, d.Y, d.Z = X, Y, Z
        d.ox, d.oy, d.oz = ox, oy, oz
        d.XYZ = XYZ if XYZ is not None else (X + 2 * ox) * (Y + 2 * oy) * (Z + 2 * oz)
        d.even_iter = 1
        d.lbm_viscosity = 0.0
        d.store_range = 1.0
        return d


class D3Q7Data(C.Structure):
    _fields_ = [("struct_size", C.c_uint32), ("version", C.c_uint32),
                ("X", C.c_int64), ("Y", C.c_int64), ("Z", C.c_int64),
                ("ox", C.c_int64), ("oy", C.c_int64), ("oz", C.c_int64), ("XYZ", C.c_int64),
                ("even_iter", C.c_int32), ("collision", C.c_int32),
                ("lbm_diffusivity", C.c_double),
                ("df", C.c_void_p * 2), ("macro", C.c_void_p), ("map", C.c_void_p),
                ("diffusion_coefficient", C.c_void_p), ("source", C.c_void_p),
                ("inflow_phi", C.c_double), ("phi_ref", C.c_double), ("store_scale", C.c_double),
                ("rows", C.c_void_p), ("row_count", C.c_int64 * 3)]

    @classmethod
    def make(cls, X, Y, Z, ox=0, oy=0, oz=0, XYZ=None):
        d = cls()
        d.struct_size = C.sizeof(cls)
        d.version = ABI_VERSION
        d.X, d.Y, d.Z = X, Y, Z
        d.ox, d.oy, d.oz = ox, oy, oz
        d.XYZ = XYZ if XYZ is not None else (X + 2 * ox) * (Y + 2 * oy) * (Z + 2 * oz)
        d.even_iter = 1
        d.phi_ref = 0.0
        d.store_scale = 1.0
        return d


I64x3 = C.c_int64 * 3
_P = C.c_void_p


def _declare(lib, prefix: str):
    """Attach argtypes/restype to the entry points shared by both libraries."""
    cfgp, d27p, d7p = C.POINTER(Config), C.POINTER(D3Q27Data), C.POINTER(D3Q7Data)
    sig = {
        "d3q27_step": [cfgp, d27p, I64x3, I64x3, C.c_uint32],
        "d3q27_stream_only": [cfgp, d27p],
        "nse_ade_step": [cfgp, d27p, d7p, I64x3, I64x3, C.c_uint32],
        "d3q27_set_equilibrium": [cfgp, d27p, C.c_double, C.c_double, C.c_double, C.c_double],
        "d3q27_set_equilibrium_field": [cfgp, d27p, _P, _P, _P, _P],
        "d3q27_compute_macro": [cfgp, d27p],
        "d3q27_export_f64": [cfgp, d27p, C.c_int, _P],
        "d3q27_import_f64": [cfgp, d27p, C.c_int, _P],
        "d3q7_set_equilibrium": [cfgp, d7p, _P, C.c_double],
        "d3q7_export_f64": [cfgp, d7p, C.c_int, _P],
        "d3q27_adjoint_step": [cfgp, d27p, _P, _P, _P],
        "d3q27_step_record": [cfgp, d27p, _P, C.c_int],
        "d3q27_adjoint_step_recorded": [cfgp, d27p, _P, C.c_int, _P, _P, _P],
        "d3q7_adjoint_step": [cfgp, d27p, d7p, _P, _P, _P, _P],
        "d3q27_halo_x": [cfgp, d27p, d27p, C.c_int],
    }
    for name, args in sig.items():
        fn = getattr(lib, prefix + name, None)
        if fn is None:
            continue
        fn.argtypes = args + ([C.c_int] if prefix == "oracle_" else [_P])
        if prefix == "oracle_" and name in ("d3q27_set_equilibrium", "d3q27_set_equilibrium_field",
                                            "d3q27_compute_macro", "d3q27_export_f64", "d3q27_import_f64",
                                            "d3q7_set_equilibrium", "d3q7_export_f64", "d3q27_adjoint_step",
                                            "d3q27_step_record", "d3q27_adjoint_step_recorded", "d3q7_adjoint_step"):
            fn.argtypes = args  # the oracle's init/transfer calls take no thread count
        fn.restype = C.c_int


PKG_DIR = Path(__file__).resolve().parent
LIB_NAME = "libtnl_lbm_amd.so"


def library_path() -> Path:
    return PKG_DIR / LIB_NAME


_lib = None


def load_library():
    """Load the HIP library; raises if it has not been built (no fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    path = library_path()
    if not path.exists():
        raise RuntimeError(f"{path} is missing: build it with `python __graft_entry__.py build` "
                           "(there is no CPU fallback for the lattice update)")
    lib = C.CDLL(str(path))
    _declare(lib, "tnl_lbm_")
    lib.tnl_lbm_abi_version.restype = C.c_uint32
    lib.tnl_lbm_last_error.restype = C.c_char_p
    lib.tnl_lbm_set_wide_addressing.argtypes = [C.c_int]
    lib.tnl_lbm_set_wide_addressing.restype = C.c_int
    lib.tnl_lbm_classify_rows.argtypes = [C.POINTER(D3Q27Data), C.POINTER(D3Q7Data), _P, C.POINTER(C.c_int64), _P]
    lib.tnl_lbm_classify_rows.restype = C.c_int
    lib.tnl_lbm_d3q27_reduce_macro.argtypes = [C.POINTER(Config), C.POINTER(D3Q27Data),
                                               C.POINTER(C.c_double), _P]
    lib.tnl_lbm_d3q27_reduce_macro.restype = C.c_int
    lib.tnl_lbm_d3q27_reduce_gradients.argtypes = [C.POINTER(Config), C.POINTER(D3Q27Data), C.c_uint32,
                                                   C.POINTER(C.c_double), _P]
    lib.tnl_lbm_d3q27_reduce_gradients.restype = C.c_int
    if lib.tnl_lbm_abi_version() != ABI_VERSION:
        raise RuntimeError("libtnl_lbm_amd.so ABI version mismatch")
    _lib = lib
    return lib


class LBMError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"tnl_lbm error {code}: {msg}")
        self.code = code


def check(lib, rc, prefix="tnl_lbm_"):
    if rc != 0:
        msg = getattr(lib, prefix + "last_error")()
        if isinstance(msg, bytes):
            msg = msg.decode()
        raise LBMError(rc, msg)
