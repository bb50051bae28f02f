"""Device lattice objects driving libtnl_lbm_amd.so through its C ABI.

PyTorch provides device memory and the stream; every lattice operation is a
call into the HIP library.  There is no CPU fallback: constructing a lattice
without the built library, or without a GPU, raises.

Storage layout is the reference's (TNL-LBM defs.h:90, lbm_data.h:50-54):
populations as a (27, X, Z, Y) tensor, i.e. index q*XYZ + (x*Z + z)*Y + y.
The `*_xyz` accessors return [x, y, z]-ordered views of it.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import torch

from . import _abi
from ._abi import GEO

_TORCH_ELEM = {_abi.STORE_F64: torch.float64, _abi.STORE_F32: torch.float32, _abi.STORE_F32_DEV: torch.float32,
               _abi.STORE_F16_DEV: torch.float16, _abi.STORE_BF16_DEV: torch.bfloat16, _abi.STORE_I16_DEV: torch.int16}


# Placement probing of the population buffer (D3Q27Lattice._allocate_populations):
# buffers of at least this size are probed under placement="auto"
PLACEMENT_MIN_BYTES = 4 << 30
# a candidate whose streaming-only step reaches this rate (0.75 of 8 TB/s) is
# kept without trying another: the A-A streaming kernels measured 0.63-0.80 of
# 8 TB/s over 768^3 FP64 buffers, from one allocation to the next and from one
# device to another
PLACEMENT_GOOD_GBS = 6000.0
# device memory left free after a further candidate
PLACEMENT_RESERVE_BYTES = 16 << 30
# spacer allocations made before the second, third, ... candidate: where a
# 98 GB buffer lands depends on what is allocated before it, and candidates
# allocated in the same state repeat the same few placements (768^3 FP64, one
# box: 63.8, 66.1, 70.0 and 63.9% in turn without spacers; after a spacer
# 63.8-78.1%, changing with the spacer's size and with what is held)
PLACEMENT_SPACERS = (6 << 30, 30 << 30, 14 << 30, 0, 2 << 30)


def _stream_ptr(stream=None):
    s = stream if stream is not None else torch.cuda.current_stream()
    return C.c_void_p(s.cuda_stream)


def _ptr(t):
    return C.c_void_p(t.data_ptr()) if t is not None else None


def i16_class_ranges(u_max: float, margin: float = 1.25):
    """Bounds of |f - w_q| / w_q per direction class |c_q|^2 = n for a flow
    with |u| <= u_max near rho = 1, for the I16 deviation form: the
    equilibrium's deviation 3 c.u + 4.5 (c.u)^2 - 1.5 u^2 with |c.u| <=
    sqrt(n) u_max, a density excursion 3 u_max^2 on top, times a margin for
    the non-equilibrium part."""
    d = 3.0 * u_max * u_max
    return tuple(margin * ((1 + d) * (3 * n ** 0.5 * u_max + 4.5 * n * u_max ** 2 + 1.5 * u_max ** 2) + d)
                 for n in range(4))


def classify_rows(nse: "D3Q27Lattice", ade: "D3Q7Lattice | None"):
    """tnl_lbm_classify_rows: the row lists of nse's map (and ade's, for the
    coupled step) as an int32 device tensor of X*Z entries and the three
    counts (classes 0, 1, 2); (None, zeros) where the lists do not apply (an
    x or z extent of 65536 or more; a coupled pair of different geometries,
    which the coupled step rejects itself): every row then steps through the
    general kernel."""
    if nse.X >= 65536 or nse.Z >= 65536:
        return None, (0, 0, 0)
    if ade is not None and (ade.d.ox, ade.d.oy, ade.d.oz, ade.d.XYZ) != (nse.d.ox, nse.d.oy, nse.d.oz, nse.d.XYZ):
        return None, (0, 0, 0)
    rows = torch.empty(nse.X * nse.Z, dtype=torch.int32, device=nse.device)
    counts = (C.c_int64 * 3)()
    nse._check(nse.lib.tnl_lbm_classify_rows(C.byref(nse.d), C.byref(ade.d) if ade is not None else None,
                                             rows.data_ptr(), counts, _stream_ptr()))
    return rows, tuple(counts)


def to_storage(a: torch.Tensor) -> torch.Tensor:
    """[x, y, z] array -> contiguous storage order (x, z, y)."""
    return a.permute(0, 2, 1).contiguous()


def from_storage(a: torch.Tensor) -> torch.Tensor:
    """storage (x, z, y) -> [x, y, z] view."""
    return a.permute(0, 2, 1)


class D3Q27Lattice:
    """One lattice block on one GPU: the device side of LBM_BLOCK
    (lbm_block.hpp:549-617) plus the step driver of State::SimUpdate
    (state.hpp:1235-1409) and LBM::updateKernelData (lbm.hpp:307-323)."""

    def __init__(self, X: int, Y: int, Z: int, *, collision="CUM", storage="f64", pattern="AA",
                 macro_real=None, nu: float = 0.0, force=(0.0, 0.0, 0.0), geo=None, device="cuda:0",
                 XYZ: int | None = None, store_range: float = 1.0, store_range_class=None, overlap=(0, 0, 0),
                 row_lists: bool = True, placement: str = "auto", placement_tries: int = 6):
        """row_lists: classify the map's rows (tnl_lbm_classify_rows) so that
        whole-box steps run the rows of FLUID/PERIODIC (and WALL) sites
        through lean kernels; False steps every row through the kernel with
        all boundary codes (same results, bit for bit).

        placement: "first" keeps the first population buffer allocated;
        "probe" times candidate buffers and keeps the fastest (up to
        placement_tries of them, see _allocate_populations); "auto" probes
        A-A lattices whose population buffer holds at least
        PLACEMENT_MIN_BYTES, and keeps the first buffer otherwise."""
        if not torch.cuda.is_available():
            raise RuntimeError("D3Q27Lattice needs a GPU (libtnl_lbm_amd.so runs on gfx950 only)")
        self.lib = _abi.load_library()
        coll = _abi.COLLISION_NAMES[collision] if isinstance(collision, str) else collision
        st = _abi.STORAGE_NAMES[storage] if isinstance(storage, str) else storage
        pat = {"AA": _abi.PATTERN_AA, "AB": _abi.PATTERN_AB}[pattern] if isinstance(pattern, str) else pattern
        if macro_real is None:
            macro_real = _abi.REAL_F64 if st == _abi.STORE_F64 else _abi.REAL_F32
        self.cfg = _abi.Config.make(coll, st, pat, macro_real)
        self.d = _abi.D3Q27Data.make(X, Y, Z, *overlap, XYZ=XYZ)
        self.ox, self.oy, self.oz = overlap
        self.d.store_range = store_range
        if store_range_class is not None:
            for n in range(4):
                self.d.store_range_class[n] = store_range_class[n]
        self.X, self.Y, self.Z = X, Y, Z
        self.device = torch.device(device)
        n = self.d.XYZ
        el = _TORCH_ELEM[st]
        ncop = 2 if pat == _abi.PATTERN_AB else 1
        self.d.lbm_viscosity = nu if nu else 1.0 / 6.0  # the probe's launches need a non-zero value
        self.df = self._allocate_populations(ncop, el, n, placement, placement_tries)
        mdt = torch.float64 if macro_real == _abi.REAL_F64 else torch.float32
        self.macro = torch.zeros(4 * n, dtype=mdt, device=self.device)
        self.map = None
        self.row_lists = row_lists
        self.rows, self._rows_key, self._geo_gen = None, None, 0
        self.d.lbm_viscosity = nu
        self.d.fx, self.d.fy, self.d.fz = force
        self._profiles = [None, None, None]
        if geo is not None:
            self.set_geometry(geo)
        self.iterations = 0
        self._bind()

    # ---- population buffers ----------------------------------------------------
    def _allocate_populations(self, ncop, el, n, placement, tries):
        """The population buffers.  Where a buffer lands in HBM changes from
        one allocation to the next and stays for the allocation's lifetime:
        the same A-A step over two buffers of one size can run at different
        rates (768^3 FP64: from 0.63 to 0.80 of 8 TB/s), and the step over a
        buffer runs within 1% of the streaming-only step over it.  A probed
        candidate is timed right after its allocation by two untimed and four
        timed streaming-only steps (its contents are undefined until the
        populations are set).  Until the fastest so far reaches
        PLACEMENT_GOOD_GBS, a further candidate is allocated while that one
        is held, right after a spacer allocation of the next size in
        PLACEMENT_SPACERS (so that it starts elsewhere; the spacer is freed
        once the candidate exists), and the slower of the two is freed and
        handed back to the driver (torch.cuda.empty_cache, so that the next
        allocation is not served from torch's cache), up to `tries`
        candidates and as long as the device has room for the spacer, one
        more candidate, the macro array and PLACEMENT_RESERVE_BYTES.
        self.placement records the probed rates in GB/s (algorithmic bytes
        of a step over its duration) and the index kept."""
        size = 27 * n * torch.tensor([], dtype=el).element_size()
        if placement not in ("auto", "probe", "first"):
            raise ValueError(f"placement {placement!r}: expected 'auto', 'probe' or 'first'")
        probe_ok = self.cfg.pattern == _abi.PATTERN_AA
        if placement == "probe" and not probe_ok:
            raise ValueError("placement probing runs the periodic A-A streaming kernels: an A-A lattice only")
        probe = probe_ok and (placement == "probe" or (placement == "auto" and size >= PLACEMENT_MIN_BYTES))
        self.placement = None
        if not probe:
            return [torch.empty(27 * n, dtype=el, device=self.device) for _ in range(ncop)]
        best = torch.empty(27 * n, dtype=el, device=self.device)
        rates = [self._probe_rate(best)]
        keep = 0
        while rates[keep] < PLACEMENT_GOOD_GBS and len(rates) < tries:
            spacer = PLACEMENT_SPACERS[(len(rates) - 1) % len(PLACEMENT_SPACERS)]
            if torch.cuda.mem_get_info(self.device)[0] < spacer + size + 4 * n * 8 + PLACEMENT_RESERVE_BYTES:
                break
            sp = torch.empty(spacer, dtype=torch.uint8, device=self.device) if spacer else None
            cand = torch.empty(27 * n, dtype=el, device=self.device)
            del sp
            rates.append(self._probe_rate(cand))
            if rates[-1] > rates[keep]:
                best, keep = cand, len(rates) - 1
            del cand
            # the spacer and the slower candidate go back to the driver, not to
            # torch's cache, from which the next allocation would be served
            torch.cuda.empty_cache()
        self.placement = {"probed_GBs": [round(r, 1) for r in rates], "kept": keep}
        return [best]

    def _probe_rate(self, buf):
        """Algorithmic GB/s of streaming-only A-A steps over `buf` (even and
        odd alternating, the local sites without the overlaps), timed with
        events on the current stream."""
        saved = self.d.df[0], self.d.df[1], self.d.even_iter
        self.d.df[0], self.d.df[1] = buf.data_ptr(), None
        s = torch.cuda.current_stream(self.device)
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
        for i in range(6):
            if i == 2:
                ev[0].record(s)
            self.d.even_iter = 1 - i % 2
            self._check(self.lib.tnl_lbm_d3q27_stream_only(C.byref(self.cfg), C.byref(self.d), _stream_ptr(s)))
        ev[1].record(s)
        ev[1].synchronize()
        self.d.df[0], self.d.df[1], self.d.even_iter = saved
        step_bytes = 2 * 27 * buf.element_size() * self.X * self.Y * self.Z
        return step_bytes * 4 / (ev[0].elapsed_time(ev[1]) * 1e-3) / 1e9

    # ---- bookkeeping --------------------------------------------------------
    @property
    def storage(self):
        return self.cfg.storage

    @property
    def nsites(self):
        return self.X * self.Y * self.Z

    def _bind(self):
        self.d.df[0] = self.df[0].data_ptr()
        self.d.df[1] = self.df[1].data_ptr() if len(self.df) > 1 else None
        self.d.macro = self.macro.data_ptr()
        self.d.map = self.map.data_ptr() if self.map is not None else None
        self.d.force_field = self.force_field.data_ptr() if getattr(self, "force_field", None) is not None else None
        key = (self._geo_gen, self.d.map, self.row_lists)
        if key != self._rows_key:
            self.rows, counts = (classify_rows(self, None) if self.row_lists and self.map is not None
                                 else (None, (0, 0, 0)))
            self._rows_key = key
            for c in range(3):
                self.d.row_count[c] = counts[c]
        self.d.rows = self.rows.data_ptr() if self.rows is not None else None

    def refresh_geometry(self):
        """After the map was changed in place: classify its rows again."""
        self._geo_gen += 1
        self._bind()

    def _check(self, rc):
        _abi.check(self.lib, rc)

    def set_geometry(self, geo):
        """geo: int16 [x, y, z] array of GEO codes over the local sites (X, Y, Z),
        or over the stored lattice with its overlaps; None = all periodic.  The
        kernels index the map as they index the populations, so a local array
        is placed inside the stored extents, the overlap sites coded PERIODIC
        (the step never reads a ghost site's own code)."""
        if geo is None:
            self.map = None
        else:
            g = torch.as_tensor(geo, dtype=torch.int16)
            Xs, Ys, Zs = self.X + 2 * self.ox, self.Y + 2 * self.oy, self.Z + 2 * self.oz
            if tuple(g.shape) == (self.X, self.Y, self.Z) and (self.ox or self.oy or self.oz):
                full = torch.full((Xs, Ys, Zs), GEO.PERIODIC, dtype=torch.int16)
                full[self.ox:self.ox + self.X, self.oy:self.oy + self.Y, self.oz:self.oz + self.Z] = g
                g = full
            if tuple(g.shape) != (Xs, Ys, Zs):
                raise ValueError(f"geometry of shape {tuple(g.shape)}: expected {(self.X, self.Y, self.Z)} or the "
                                 f"stored {(Xs, Ys, Zs)}")
            self.map = to_storage(g).reshape(-1).to(self.device)
        self._geo_gen += 1
        if hasattr(self, "iterations"):
            self._bind()

    def set_force_field(self, fx=None, fy=None, fz=None):
        """MACRO::computeForcing's per-site force (macro.h:50-52, called at
        kernels.h:103), added to the homogeneous force before each collision:
        [x, y, z] arrays over the stored lattice (overlaps included), None for
        a zero component; all three None removes the field.  Kept as three
        planes of the macro element type in the lattice layout."""
        if fx is None and fy is None and fz is None:
            self.force_field = None
        else:
            mdt = torch.float64 if self.cfg.macro_real == _abi.REAL_F64 else torch.float32
            Xs, Ys, Zs = self.X + 2 * self.ox, self.Y + 2 * self.oy, self.Z + 2 * self.oz
            t = torch.zeros(3, self.d.XYZ, dtype=mdt, device=self.device)
            for k, a in enumerate((fx, fy, fz)):
                if a is not None:
                    a = torch.as_tensor(a, dtype=mdt)
                    assert tuple(a.shape) == (Xs, Ys, Zs), (tuple(a.shape), (Xs, Ys, Zs))
                    t[k, :Xs * Ys * Zs] = to_storage(a).reshape(-1).to(self.device)
            self.force_field = t
        self._bind()

    def set_inflow(self, vx=0.0, vy=0.0, vz=0.0, profile=None, profile_real=None):
        """NSE_Data_ConstInflow values, or NSE_Data_InflowProfile arrays of shape (Z, Y).
        profile_real: element type of the profile arrays, "f64" (default) or
        "f32" (the reference's TraitsSP profiles hold float, lbm_data.h:100-104)."""
        self.d.inflow_vx, self.d.inflow_vy, self.d.inflow_vz = vx, vy, vz
        self._profiles = [None, None, None]
        f32 = profile_real == "f32"
        if profile_real not in (None, "f64", "f32"):
            raise ValueError(f"profile_real must be 'f64' or 'f32', not {profile_real!r}")
        self.d.inflow_profile_real = _abi.REAL_F32 if f32 else _abi.REAL_F64
        for k in range(3):
            if profile is not None and profile[k] is not None:
                p = torch.as_tensor(profile[k], dtype=torch.float32 if f32 else torch.float64,
                                    device=self.device).contiguous()
                assert p.numel() == self.Y * self.Z
                self._profiles[k] = p
                self.d.inflow_profile[k] = p.data_ptr()
            else:
                self.d.inflow_profile[k] = None

    # ---- initialisation -----------------------------------------------------
    def set_equilibrium(self, rho=1.0, vx=0.0, vy=0.0, vz=0.0):
        self._check(self.lib.tnl_lbm_d3q27_set_equilibrium(C.byref(self.cfg), C.byref(self.d), rho, vx, vy, vz,
                                                           _stream_ptr()))

    def set_equilibrium_field(self, rho, vx, vy, vz):
        """Fields are [x, y, z] float64 arrays (host or device)."""
        fs = [to_storage(torch.as_tensor(a, dtype=torch.float64)).reshape(-1).to(self.device) for a in (rho, vx, vy, vz)]
        if self.d.XYZ != self.nsites:
            raise ValueError("field initialisation expects XYZ == X*Y*Z")
        self._check(self.lib.tnl_lbm_d3q27_set_equilibrium_field(C.byref(self.cfg), C.byref(self.d),
                                                                 *[_ptr(f) for f in fs], _stream_ptr()))
        torch.cuda.current_stream().synchronize()

    def compute_macro(self):
        self._check(self.lib.tnl_lbm_d3q27_compute_macro(C.byref(self.cfg), C.byref(self.d), _stream_ptr()))

    # ---- the step -------------------------------------------------------------
    def update_kernel_data(self):
        """LBM::updateKernelData (lbm.hpp:307-323): A-A parity and A-B rotation."""
        self.d.even_iter = 1 if self.iterations % 2 == 0 else 0
        if self.cfg.pattern == _abi.PATTERN_AB:
            cur = self.iterations % 2
            self.d.df[0] = self.df[cur].data_ptr()
            self.d.df[1] = self.df[1 - cur].data_ptr()

    def step(self, output_macro=False, stream=None, flags=0, offset=(0, 0, 0), end=None):
        self.update_kernel_data()
        end = end or (self.X, self.Y, self.Z)
        fl = flags | (_abi.FLAG_OUTPUT_MACRO if output_macro else 0)
        self._check(self.lib.tnl_lbm_d3q27_step(C.byref(self.cfg), C.byref(self.d), _abi.I64x3(*offset),
                                                _abi.I64x3(*end), fl, _stream_ptr(stream)))
        self.iterations += 1

    def stream_only(self, stream=None):
        """One A-A step of free streaming (the collision replaced by the
        identity) through the periodic step kernels: a measurement of the
        step's memory ceiling on this allocation (tnl_lbm_d3q27_stream_only)."""
        self.update_kernel_data()
        self._check(self.lib.tnl_lbm_d3q27_stream_only(C.byref(self.cfg), C.byref(self.d), _stream_ptr(stream)))
        self.iterations += 1

    def step_coupled(self, ade: "D3Q7Lattice", output_macro=False, stream=None, flags=0, offset=(0, 0, 0), end=None):
        """One cudaLBMKernel<NSE,ADE> launch (kernels.h:114-196, state_NSE_ADE.h:201-210): the D3Q27
        update and, at the same site, the D3Q7 update driven by the NSE velocity."""
        if ade.nse is not self:
            raise ValueError("the D3Q7 lattice belongs to another NSE lattice")
        ade.refresh_rows()
        self.update_kernel_data()
        ade.update_kernel_data(self.iterations)
        end = end or (self.X, self.Y, self.Z)
        fl = flags | (_abi.FLAG_OUTPUT_MACRO if output_macro else 0)
        self._check(self.lib.tnl_lbm_nse_ade_step(C.byref(self.cfg), C.byref(self.d), C.byref(ade.d),
                                                  _abi.I64x3(*offset), _abi.I64x3(*end), fl, _stream_ptr(stream)))
        self.iterations += 1

    def current_df(self):
        return self.iterations % 2 if self.cfg.pattern == _abi.PATTERN_AB else 0

    def synchronize(self):
        """Wait for the work queued on this lattice's device."""
        torch.cuda.synchronize(self.device)

    def halo_x(self, upper: "D3Q27Lattice", phase: int, stream=None):
        """Halo exchange of an x-slab split (tnl_lbm_d3q27_halo_x): this slab's
        +x face meets `upper`'s -x face (`upper` is this slab for a slab that
        is its own periodic neighbour).  phase _abi.HALO_PULL before an A-B
        step or after an A-A even step, _abi.HALO_PUSH after an A-A odd step.
        Both slabs need overlap (1, .., ..) or more and step with
        _abi.FLAG_DISTRIBUTED_X."""
        for lat in (self, upper):
            lat.update_kernel_data()  # df[0]: the array the next step reads
        self._check(self.lib.tnl_lbm_d3q27_halo_x(C.byref(self.cfg), C.byref(self.d), C.byref(upper.d), phase,
                                                  _stream_ptr(stream)))

    def stored(self, which: int = 0) -> torch.Tensor:
        """Population array `which` as a (27, X + 2ox, Z + 2oz, Y + 2oy) view of
        the stored layout (q slowest, y fastest), overlaps included."""
        return self.df[which].view(27, self.X + 2 * self.ox, self.Z + 2 * self.oz, self.Y + 2 * self.oy)

    # ---- adjoint ----------------------------------------------------------------
    def adjoint_step(self, lam_next: torch.Tensor, lam: torch.Tensor, grad_inflow: torch.Tensor | None = None,
                     which: int | None = None, stream=None):
        """lam <- J^T lam_next for one A-B step taken from df[which] (default: the
        current state); grad_inflow (3, Z, Y) float64 accumulates the gradient
        with respect to the inflow velocity.  lam_next / lam are flat 27*XYZ
        tensors of the storage dtype in the population layout."""
        which = self.current_df() if which is None else which
        el = _TORCH_ELEM[self.cfg.storage]
        for t in (lam_next, lam):
            assert t.dtype == el and t.numel() == 27 * self.d.XYZ and t.is_contiguous()
        if grad_inflow is not None:
            assert grad_inflow.dtype == torch.float64 and grad_inflow.numel() == 3 * self.Y * self.Z
        saved = self.d.df[0], self.d.df[1]
        self.d.df[0] = self.df[which].data_ptr()
        try:
            self._check(self.lib.tnl_lbm_d3q27_adjoint_step(C.byref(self.cfg), C.byref(self.d), _ptr(lam_next), _ptr(lam),
                                                            _ptr(grad_inflow), _stream_ptr(stream)))
        finally:
            self.d.df[0], self.d.df[1] = saved

    # ---- readback -------------------------------------------------------------
    def export_f64(self, which=None) -> torch.Tensor:
        """Absolute populations, slot by slot, as a (27, X, Y, Z) float64 view."""
        which = self.current_df() if which is None else which
        out = torch.empty(27 * self.d.XYZ, dtype=torch.float64, device=self.device)
        saved = self.d.df[0], self.d.df[1]
        self.d.df[0] = self.df[which].data_ptr()
        try:
            self._check(self.lib.tnl_lbm_d3q27_export_f64(C.byref(self.cfg), C.byref(self.d), 0, _ptr(out),
                                                          _stream_ptr()))
        finally:
            self.d.df[0], self.d.df[1] = saved
        return out.view(27, self.d.XYZ)[:, :self.nsites].reshape(27, self.X, self.Z, self.Y).permute(0, 1, 3, 2)

    def import_f64(self, f, which=0):
        f = torch.as_tensor(f, dtype=torch.float64, device=self.device).permute(0, 1, 3, 2).contiguous()
        if self.d.XYZ != self.nsites:
            full = torch.zeros(27, self.d.XYZ, dtype=torch.float64, device=self.device)
            full[:, :self.nsites] = f.reshape(27, -1)
            f = full
        f = f.reshape(-1)
        self._check(self.lib.tnl_lbm_d3q27_import_f64(C.byref(self.cfg), C.byref(self.d), which, _ptr(f),
                                                      _stream_ptr()))
        torch.cuda.current_stream().synchronize()

    def macros(self):
        """rho, vx, vy, vz as [x, y, z] views of the macro array."""
        m = self.macro.view(4, self.d.XYZ)[:, :self.nsites].view(4, self.X, self.Z, self.Y).permute(0, 1, 3, 2)
        return m[0], m[1], m[2], m[3]

    def reduce_macro(self):
        """FP64 sums: mass, sum rho|u|^2, NaN count, sum (rho-1)^2, sum |u|^2, n."""
        out = (C.c_double * 6)()
        self._check(self.lib.tnl_lbm_d3q27_reduce_macro(C.byref(self.cfg), C.byref(self.d), out, _stream_ptr()))
        return list(out)

    def reduce_gradients(self, flags=0):
        """FP64 sums of rho |grad u|^2 and rho |curl u|^2 over the macro array with
        the central differences of sim_4.cu's computeGradVel (sim_4.cu:20-82):
        the enstrophy diagnostic before its 1/(2 rho_0 V) normalisation."""
        out = (C.c_double * 2)()
        self._check(self.lib.tnl_lbm_d3q27_reduce_gradients(C.byref(self.cfg), C.byref(self.d), flags, out,
                                                            _stream_ptr()))
        return list(out)


class D3Q7Lattice:
    """The advection-diffusion block of State_NSE_ADE (state_NSE_ADE.h:3-401):
    D3Q7 populations in the storage form and streaming pattern of its NSE
    lattice, stepped together with it by D3Q27Lattice.step_coupled.

    ADE_Data (lbm_data.h:212-261): constant or per-site diffusivity D (the
    relaxation time is 1/2 + 4D), D3Q7 geometry codes (ADE_GEO), an inflow
    value, and a per-site source S added as w_q S after the collision.  The
    deviation storage forms store (f_q - w_q phi_ref) * store_scale."""

    def __init__(self, nse: D3Q27Lattice, *, collision="SRT", diffusivity: float = 0.0, geo=None,
                 phi_ref: float = 0.0, store_scale: float = 1.0):
        self.nse = nse
        self.lib = nse.lib
        self.cfg = nse.cfg
        self.X, self.Y, self.Z = nse.X, nse.Y, nse.Z
        self.device = nse.device
        self.d = _abi.D3Q7Data.make(nse.X, nse.Y, nse.Z, XYZ=nse.d.XYZ)
        self.d.collision = _abi.ADE_COLLISION_NAMES[collision] if isinstance(collision, str) else collision
        self.d.lbm_diffusivity = diffusivity
        self.d.phi_ref, self.d.store_scale = phi_ref, store_scale
        n = self.d.XYZ
        el = _TORCH_ELEM[self.cfg.storage]
        ncop = 2 if self.cfg.pattern == _abi.PATTERN_AB else 1
        self.df = [torch.empty(7 * n, dtype=el, device=self.device) for _ in range(ncop)]
        self._mdt = torch.float64 if self.cfg.macro_real == _abi.REAL_F64 else torch.float32
        self.phi = torch.zeros(n, dtype=self._mdt, device=self.device)
        self.map = self.diffusion = self.source = None
        self.rows, self._rows_key, self._geo_gen = None, None, 0
        if geo is not None:
            self.set_geometry(geo)
        self._bind()

    @property
    def nsites(self):
        return self.X * self.Y * self.Z

    def _bind(self):
        self.d.df[0] = self.df[0].data_ptr()
        self.d.df[1] = self.df[1].data_ptr() if len(self.df) > 1 else None
        self.d.macro = self.phi.data_ptr()
        self.d.map = self.map.data_ptr() if self.map is not None else None
        self.d.diffusion_coefficient = self.diffusion.data_ptr() if self.diffusion is not None else None
        self.d.source = self.source.data_ptr() if self.source is not None else None
        self.refresh_rows()

    def refresh_rows(self):
        """The coupled step's row lists, classified from this map and the NSE
        lattice's, again where either has changed since."""
        nse = self.nse
        key = (nse._geo_gen, nse.d.map, self._geo_gen, self.d.map, nse.row_lists)
        if key != self._rows_key:
            self.rows, counts = classify_rows(nse, self) if nse.row_lists else (None, (0, 0, 0))
            self._rows_key = key
            for c in range(3):
                self.d.row_count[c] = counts[c]
        self.d.rows = self.rows.data_ptr() if self.rows is not None else None

    def refresh_geometry(self):
        """After the map was changed in place: classify the rows again."""
        self._geo_gen += 1
        self._bind()

    def _field(self, a):
        """[x, y, z] array -> flat device array in the lattice layout (macro_real)."""
        t = torch.as_tensor(a, dtype=self._mdt)
        assert tuple(t.shape) == (self.X, self.Y, self.Z)
        flat = torch.zeros(self.d.XYZ, dtype=self._mdt, device=self.device)
        flat[:self.nsites] = to_storage(t).reshape(-1).to(self.device)
        return flat

    def set_geometry(self, geo):
        """geo: int16 [x, y, z] array of D3Q7 GEO codes (None = all periodic)."""
        if geo is None:
            self.map = None
        else:
            g = torch.as_tensor(geo, dtype=torch.int16)
            assert tuple(g.shape) == (self.X, self.Y, self.Z)
            self.map = to_storage(g).reshape(-1).to(self.device)
        self._geo_gen += 1
        self._bind()

    def set_diffusivity_field(self, D):
        """Per-site diffusivity (ADE_Data::diffusionCoefficient, lbm_data.h:231-239); None = constant."""
        self.diffusion = None if D is None else self._field(D)
        self._bind()

    def set_source(self, S):
        """Per-site source S: every collided site gains w_q S (None = no source)."""
        self.source = None if S is None else self._field(S)
        self._bind()

    def set_inflow(self, phi):
        self.d.inflow_phi = phi

    def set_equilibrium(self, phi=0.0):
        """f_q = w_q phi (D3Q7 equilibrium at rest) at every stored site; phi is a scalar or an [x, y, z] field."""
        field = None if isinstance(phi, (int, float)) else self._field(phi)
        self._check(self.lib.tnl_lbm_d3q7_set_equilibrium(C.byref(self.cfg), C.byref(self.d), _ptr(field),
                                                          float(phi) if field is None else 0.0, _stream_ptr()))
        torch.cuda.current_stream().synchronize()

    def update_kernel_data(self, iterations):
        self.d.even_iter = 1 if iterations % 2 == 0 else 0
        if self.cfg.pattern == _abi.PATTERN_AB:
            cur = iterations % 2
            self.d.df[0] = self.df[cur].data_ptr()
            self.d.df[1] = self.df[1 - cur].data_ptr()

    def _check(self, rc):
        _abi.check(self.lib, rc)

    def export_f64(self, which=None) -> torch.Tensor:
        """Absolute populations, slot by slot, as a (7, X, Y, Z) float64 view."""
        which = self.nse.current_df() if which is None else which
        out = torch.empty(7 * self.d.XYZ, dtype=torch.float64, device=self.device)
        saved = self.d.df[0], self.d.df[1]
        self.d.df[0] = self.df[which].data_ptr()
        try:
            self._check(self.lib.tnl_lbm_d3q7_export_f64(C.byref(self.cfg), C.byref(self.d), 0, _ptr(out),
                                                         _stream_ptr()))
        finally:
            self.d.df[0], self.d.df[1] = saved
        return out.view(7, self.d.XYZ)[:, :self.nsites].reshape(7, self.X, self.Z, self.Y).permute(0, 1, 3, 2)

    def phi_xyz(self):
        """The scalar written on output steps, as an [x, y, z] view."""
        return self.phi[:self.nsites].view(self.X, self.Z, self.Y).permute(0, 2, 1)
