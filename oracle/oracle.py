"""Python handle on the CPU oracle (liblbm_oracle.so) -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
import this module, as the checker.  The product path (pytnl_lbm over
libtnl_lbm_amd.so) never imports it.

`HostLattice` mirrors the device lattice of pytnl_lbm with numpy arrays and
calls the oracle's restatement of the reference's per-site kernel
(TNL-LBM include/lbm3d/kernels.h:59-196).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent
sys.path.insert(0, str(ROOT / "tnl-lbm_amd"))
from pytnl_lbm import _abi  # noqa: E402

LIB_PATH = HERE / "liblbm_oracle.so"
_lib = None


def build(force: bool = False) -> Path:
    if force or not LIB_PATH.exists():
        subprocess.run(["make", "-s", "-B" if force else "-s", "-C", str(HERE)], check=True)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        L = C.CDLL(str(LIB_PATH))
        _abi._declare(L, "oracle_")
        L.oracle_last_error.restype = C.c_char_p
        dp = C.POINTER(C.c_double)
        L.oracle_d3q27_collide.argtypes = [C.c_int, C.c_int, dp, C.c_double, C.c_double, C.c_double, C.c_double, dp]
        L.oracle_d3q27_collide_dev.argtypes = [C.c_int, dp, C.c_double, C.c_double, C.c_double, C.c_double, C.c_int]
        L.oracle_d3q27_equilibrium.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double, dp]
        L.oracle_d3q27_density_velocity.argtypes = [C.c_int, dp, C.c_double, C.c_double, C.c_double, dp]
        L.oracle_d3q27_adjoint_step.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        L.oracle_d3q27_adjoint_step.restype = C.c_int
        L.oracle_f32_to_f16.argtypes = [C.c_float]
        L.oracle_f32_to_f16.restype = C.c_uint16
        L.oracle_f16_to_f32.argtypes = [C.c_uint16]
        L.oracle_f16_to_f32.restype = C.c_float
        L.oracle_f32_to_bf16.argtypes = [C.c_float]
        L.oracle_f32_to_bf16.restype = C.c_uint16
        L.oracle_bf16_to_f32.argtypes = [C.c_uint16]
        L.oracle_bf16_to_f32.restype = C.c_float
        _lib = L
    return _lib


def _chk(rc):
    if rc != 0:
        raise _abi.LBMError(rc, lib().oracle_last_error().decode())


def _ptr(a):
    return C.c_void_p(a.ctypes.data) if a is not None else None


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


# ---- single-site helpers ---------------------------------------------------
def collide(f, nu, collision=_abi.COLL_CUM, F=(0.0, 0.0, 0.0), single=False):
    """One collision of one site (COLL::collision after computeDensityAndVelocity)."""
    f = np.array(f, dtype=np.float64).copy()
    out = np.zeros(4)
    _chk(lib().oracle_d3q27_collide(collision, int(single), _dp(f), nu, F[0], F[1], F[2], _dp(out)))
    return f, out


def collide_dev(g, nu, F=(0.0, 0.0, 0.0), single=True):
    g = np.array(g, dtype=np.float64).copy()
    _chk(lib().oracle_d3q27_collide_dev(_abi.COLL_CUM, _dp(g), nu, F[0], F[1], F[2], int(single)))
    return g


def equilibrium(rho, vx, vy, vz, kind=0, single=False):
    """kind 0: D3Q27_EQ_INV_CUM, kind 1: D3Q27_EQ."""
    f = np.zeros(27)
    _chk(lib().oracle_d3q27_equilibrium(kind, int(single), rho, vx, vy, vz, _dp(f)))
    return f


def density_velocity(f, F=(0.0, 0.0, 0.0), single=False):
    f = np.ascontiguousarray(f, dtype=np.float64)
    out = np.zeros(4)
    _chk(lib().oracle_d3q27_density_velocity(int(single), _dp(f), F[0], F[1], F[2], _dp(out)))
    return out


# ---- whole lattice -----------------------------------------------------------
_NP_ELEM = {_abi.STORE_F64: np.float64, _abi.STORE_F32: np.float32, _abi.STORE_F32_DEV: np.float32,
            _abi.STORE_F16_DEV: np.uint16, _abi.STORE_BF16_DEV: np.uint16, _abi.STORE_I16_DEV: np.int16}


class HostLattice:
    """A D3Q27 lattice in host memory driven by the oracle."""

    def __init__(self, X, Y, Z, *, collision=_abi.COLL_CUM, storage=_abi.STORE_F64, pattern=_abi.PATTERN_AA,
                 macro_real=_abi.REAL_F64, nu=0.0, force=(0.0, 0.0, 0.0), geo=None, threads=None,
                 store_range=1.0, store_range_class=None):
        self.cfg = _abi.Config.make(collision, storage, pattern, macro_real)
        self.d = _abi.D3Q27Data.make(X, Y, Z)
        self.d.store_range = store_range
        if store_range_class is not None:
            for n in range(4):
                self.d.store_range_class[n] = store_range_class[n]
        self.X, self.Y, self.Z = X, Y, Z
        n = self.d.XYZ
        el = _NP_ELEM[storage]
        self.df = [np.zeros(27 * n, dtype=el), np.zeros(27 * n, dtype=el) if pattern == _abi.PATTERN_AB else None]
        self.macro = np.zeros(4 * n, dtype=np.float64 if macro_real == _abi.REAL_F64 else np.float32)
        self.map = None if geo is None else np.ascontiguousarray(geo.transpose(0, 2, 1).reshape(-1), dtype=np.int16)
        self._geo_src = geo
        self.d.df[0] = self.df[0].ctypes.data
        self.d.df[1] = self.df[1].ctypes.data if self.df[1] is not None else None
        self.d.macro = self.macro.ctypes.data
        self.d.map = self.map.ctypes.data if self.map is not None else None
        self.d.lbm_viscosity = nu
        self.d.fx, self.d.fy, self.d.fz = force
        self.iterations = 0
        self.threads = threads or int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        self._profiles = []

    # layout helpers: arrays indexed [x, y, z] <-> storage (x*Z + z)*Y + y
    def to_storage(self, a):
        return np.ascontiguousarray(np.asarray(a, dtype=np.float64).transpose(0, 2, 1)).reshape(-1)

    def from_storage(self, v):
        return v.reshape(self.X, self.Z, self.Y).transpose(0, 2, 1)

    def set_force_field(self, fx=None, fy=None, fz=None):
        """The per-site force of MACRO::computeForcing: [x, y, z] arrays or None."""
        if fx is None and fy is None and fz is None:
            self._force = None
            self.d.force_field = None
            return
        dt = np.float64 if self.cfg.macro_real == _abi.REAL_F64 else np.float32
        t = np.zeros((3, self.d.XYZ), dtype=dt)
        n = self.X * self.Y * self.Z
        for k, a in enumerate((fx, fy, fz)):
            if a is not None:
                t[k, :n] = self.to_storage(a).astype(dt)
        self._force = t
        self.d.force_field = t.ctypes.data

    def set_inflow(self, vx=0.0, vy=0.0, vz=0.0, profile=None, profile_real=None):
        self.d.inflow_vx, self.d.inflow_vy, self.d.inflow_vz = vx, vy, vz
        self._profiles = []
        f32 = profile_real == "f32"
        self.d.inflow_profile_real = _abi.REAL_F32 if f32 else _abi.REAL_F64
        for k in range(3):
            if profile is not None and profile[k] is not None:
                # indexed Y*z + y: shape (Z, Y)
                p = np.ascontiguousarray(profile[k], dtype=np.float32 if f32 else np.float64)
                self._profiles.append(p)
                self.d.inflow_profile[k] = p.ctypes.data
            else:
                self.d.inflow_profile[k] = None

    def set_equilibrium_field(self, rho, vx, vy, vz):
        arrs = [self.to_storage(a) for a in (rho, vx, vy, vz)]
        _chk(lib().oracle_d3q27_set_equilibrium_field(C.byref(self.cfg), C.byref(self.d), *[_ptr(a) for a in arrs]))

    def set_equilibrium(self, rho=1.0, vx=0.0, vy=0.0, vz=0.0):
        _chk(lib().oracle_d3q27_set_equilibrium(C.byref(self.cfg), C.byref(self.d), rho, vx, vy, vz))

    def compute_macro(self):
        _chk(lib().oracle_d3q27_compute_macro(C.byref(self.cfg), C.byref(self.d)))

    def step(self, output_macro=False, flags=0):
        """One State::SimUpdate: parity and A-B rotation (lbm.hpp:307-323), then the kernel."""
        self.d.even_iter = 1 if self.iterations % 2 == 0 else 0
        if self.cfg.pattern == _abi.PATTERN_AB:
            cur = self.iterations % 2
            self.d.df[0] = self.df[cur].ctypes.data
            self.d.df[1] = self.df[1 - cur].ctypes.data
        off = _abi.I64x3(0, 0, 0)
        end = _abi.I64x3(self.X, self.Y, self.Z)
        fl = flags | (_abi.FLAG_OUTPUT_MACRO if output_macro else 0)
        _chk(lib().oracle_d3q27_step(C.byref(self.cfg), C.byref(self.d), off, end, fl, self.threads))
        self.iterations += 1

    def current_df(self):
        """Index of the array holding the latest populations."""
        if self.cfg.pattern == _abi.PATTERN_AB:
            return self.iterations % 2
        return 0

    def export_f64(self, which=None):
        which = self.current_df() if which is None else which
        out = np.zeros(27 * self.d.XYZ)
        saved = self.d.df[0], self.d.df[1]
        self.d.df[0] = self.df[which].ctypes.data
        _chk(lib().oracle_d3q27_export_f64(C.byref(self.cfg), C.byref(self.d), 0, _ptr(out)))
        self.d.df[0], self.d.df[1] = saved
        return out.reshape(27, self.X, self.Z, self.Y).transpose(0, 1, 3, 2)

    def import_f64(self, f, which=0):
        f = np.ascontiguousarray(np.asarray(f, dtype=np.float64).transpose(0, 1, 3, 2)).reshape(-1)
        _chk(lib().oracle_d3q27_import_f64(C.byref(self.cfg), C.byref(self.d), which, _ptr(f)))

    def macros(self):
        """rho, vx, vy, vz as [x, y, z] arrays (float64)."""
        m = self.macro.astype(np.float64).reshape(4, self.X, self.Z, self.Y).transpose(0, 1, 3, 2)
        return m[0], m[1], m[2], m[3]

    def adjoint_step(self, lam_next, grad=None, which=None):
        """Adjoint of one A-B step taken from the state in df[which] (default:
        the current one): returns J^T lam_next as a (27, X, Y, Z) array and adds
        d/d(inflow velocity) to grad (shape (3, Z, Y)) when given."""
        which = self.current_df() if which is None else which
        el = _NP_ELEM[self.cfg.storage]
        ln = np.ascontiguousarray(np.asarray(lam_next, dtype=np.float64).transpose(0, 1, 3, 2)).reshape(-1).astype(el)
        out = np.zeros_like(ln)
        saved = self.d.df[0]
        self.d.df[0] = self.df[which].ctypes.data
        g = None if grad is None else np.ascontiguousarray(grad, dtype=np.float64)
        try:
            _chk(lib().oracle_d3q27_adjoint_step(C.byref(self.cfg), C.byref(self.d), _ptr(ln), _ptr(out), _ptr(g)))
        finally:
            self.d.df[0] = saved
        if grad is not None:
            grad[...] = g
        return out.astype(np.float64).reshape(27, self.X, self.Z, self.Y).transpose(0, 1, 3, 2)

    def step_coupled(self, ade, output_macro=False, flags=0):
        """One coupled NSE + ADE step (kernels.h:114-196): parity and A-B rotation for both lattices."""
        self.d.even_iter = 1 if self.iterations % 2 == 0 else 0
        ade.d.even_iter = self.d.even_iter
        if self.cfg.pattern == _abi.PATTERN_AB:
            cur = self.iterations % 2
            self.d.df[0] = self.df[cur].ctypes.data
            self.d.df[1] = self.df[1 - cur].ctypes.data
            ade.d.df[0] = ade.df[cur].ctypes.data
            ade.d.df[1] = ade.df[1 - cur].ctypes.data
        off = _abi.I64x3(0, 0, 0)
        end = _abi.I64x3(self.X, self.Y, self.Z)
        fl = flags | (_abi.FLAG_OUTPUT_MACRO if output_macro else 0)
        _chk(lib().oracle_nse_ade_step(C.byref(self.cfg), C.byref(self.d), C.byref(ade.d), off, end, fl, self.threads))
        self.iterations += 1


class HostAde:
    """D3Q7 ADE populations in host memory, stepped with their HostLattice."""

    def __init__(self, nse: HostLattice, *, collision=_abi.ADE_COLL_SRT, diffusivity=0.0, geo=None, phi_ref=0.0,
                 store_scale=1.0):
        self.nse = nse
        self.cfg = nse.cfg
        self.X, self.Y, self.Z = nse.X, nse.Y, nse.Z
        self.d = _abi.D3Q7Data.make(nse.X, nse.Y, nse.Z, XYZ=nse.d.XYZ)
        self.d.collision = collision
        self.d.lbm_diffusivity = diffusivity
        self.d.phi_ref, self.d.store_scale = phi_ref, store_scale
        n = self.d.XYZ
        el = _NP_ELEM[self.cfg.storage]
        self.df = [np.zeros(7 * n, dtype=el) for _ in range(2 if self.cfg.pattern == _abi.PATTERN_AB else 1)]
        self._mdt = np.float64 if self.cfg.macro_real == _abi.REAL_F64 else np.float32
        self.phi = np.zeros(n, dtype=self._mdt)
        self.map = self.diffusion = self.source = None
        self.d.df[0] = self.df[0].ctypes.data
        self.d.df[1] = self.df[1].ctypes.data if len(self.df) > 1 else None
        self.d.macro = self.phi.ctypes.data
        if geo is not None:
            self.set_geometry(geo)

    def _field(self, a):
        return np.ascontiguousarray(np.asarray(a, dtype=np.float64).transpose(0, 2, 1), dtype=self._mdt).reshape(-1)

    def set_geometry(self, geo):
        self.map = np.ascontiguousarray(np.asarray(geo).transpose(0, 2, 1).reshape(-1), dtype=np.int16)
        self.d.map = self.map.ctypes.data

    def set_diffusivity_field(self, D):
        self.diffusion = self._field(D)
        self.d.diffusion_coefficient = self.diffusion.ctypes.data

    def set_source(self, S):
        self.source = self._field(S)
        self.d.source = self.source.ctypes.data

    def set_inflow(self, phi):
        self.d.inflow_phi = phi

    def set_equilibrium(self, phi=0.0):
        field = None if np.isscalar(phi) else self._field(phi)
        _chk(lib().oracle_d3q7_set_equilibrium(C.byref(self.cfg), C.byref(self.d), _ptr(field),
                                               float(phi) if field is None else 0.0))

    def export_f64(self, which=None):
        which = self.nse.current_df() if which is None else which
        out = np.zeros(7 * self.d.XYZ)
        saved = self.d.df[0], self.d.df[1]
        self.d.df[0] = self.df[which].ctypes.data
        try:
            _chk(lib().oracle_d3q7_export_f64(C.byref(self.cfg), C.byref(self.d), 0, _ptr(out)))
        finally:
            self.d.df[0], self.d.df[1] = saved
        return out.reshape(7, self.X, self.Z, self.Y).transpose(0, 1, 3, 2)

    def phi_xyz(self):
        return self.phi.astype(np.float64).reshape(self.X, self.Z, self.Y).transpose(0, 2, 1)
