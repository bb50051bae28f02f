"""Differentiable lattice runs: T A-B steps behind torch.autograd.Function,
with checkpointed recomputation over time.

The reference's adjoint workflow (sim_adjoint/sim_adjoint.cu:595-777) runs a
primary simulation with an inflow profile, then an adjoint simulation whose
kernel (col_srt_adjoint.h, streaming_AB.h, bc.h) accumulates the gradient of
a velocity-mismatch loss with respect to that profile.  Here the adjoint is the
exact discrete adjoint of the product's own step (tnl_lbm_d3q27_adjoint_step:
default cumulant, A-B pattern), so a loss written in torch over the final
populations differentiates back to the inflow profile and the initial state.

Memory: sqrt(T) checkpoints of the state are kept during the forward run; the
backward pass recomputes one segment of about sqrt(T) states from its
checkpoint, then runs the adjoint steps over it in reverse.  The recomputed
states are written in place (each step's output buffer is the next state's
storage), so neither the checkpoints nor the segment cost an extra copy.
"""
from __future__ import annotations

import ctypes as C
import math

import torch

from . import _abi
from .lattice import D3Q27Lattice, _ptr, _stream_ptr

# D3Q27 velocity set in the reference's direction order (defs.h:317-350)
D3Q27_CX = (0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1)
D3Q27_CY = (0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, -1, 1, -1, 1)
D3Q27_CZ = (0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, -1, 1)


def populations(f: torch.Tensor, lat: D3Q27Lattice) -> torch.Tensor:
    """Flat storage-layout populations -> (27, X, Z, Y) view."""
    return f.view(27, lat.d.XYZ)[:, :lat.nsites].view(27, lat.X, lat.Z, lat.Y)


def density_velocity(f: torch.Tensor, force=(0.0, 0.0, 0.0)):
    """rho and u = (j + F/2)/rho of populations shaped (27, ...), in FP64 and
    differentiable (D3Q27_COMMON::computeDensityAndVelocity, common.h:17-50,
    with the half-force shift of col_cum.h:20-22)."""
    f = f.double()
    c = torch.tensor([D3Q27_CX, D3Q27_CY, D3Q27_CZ], dtype=torch.float64, device=f.device)
    rho = f.sum(0)
    j = torch.tensordot(c, f, dims=([1], [0]))
    F = torch.tensor(force, dtype=torch.float64, device=f.device).view(3, *([1] * (f.dim() - 1)))
    return rho, (j + 0.5 * F) / rho


MODES = ("recompute", "segment", "tape", "hybrid", "mixed")


def tape_elements(XYZ: int) -> int:
    """Elements of one step's moment recording: blocks of 64 sites x 10 moments."""
    return (XYZ + 63) // 64 * 640


def tape_moments(tape: torch.Tensor, XYZ: int) -> torch.Tensor:
    """One step's recording -> (10, XYZ) view order (m0, jx, jy, jz, Pxx, Pyy,
    Pzz, Pxy, Pxz, Pyz)."""
    nb = (XYZ + 63) // 64
    return tape.view(nb, 10, 64).permute(1, 0, 2).reshape(10, nb * 64)[:, :XYZ]


class CheckpointedRun:
    """T A-B steps of a lattice from a given initial state, differentiable
    through the hand-written adjoint.

    The Jacobian of a step at a site depends on the pulled populations only
    through their ten raw moments, so the backward pass needs either the
    states or a recording of those moments.  Three schedules:

    - ``"segment"`` (default): a checkpoint every K steps (K = ceil(sqrt(T)));
      backwards, each segment is recomputed from its checkpoint by steps that
      record the moments (tnl_lbm_d3q27_step_record), and the adjoint steps
      read that recording.  Memory: T/K states + K recordings of 10 values.
    - ``"recompute"``: the same checkpoints; the segment's states themselves
      are kept and the adjoint pulls from them.  Memory: T/K + K states.
    - ``"tape"``: every forward step records its moments; the backward pass
      reads them and recomputes nothing.  Memory: T recordings (float32 by
      default: 40 bytes per site and step).
    - ``"hybrid"``: the checkpoints of ``"segment"``, and the forward pass
      also records the moments of as many trailing segments as
      ``record_budget`` bytes hold (default: 80% of the device memory free
      once the other buffers exist); only the segments before those are
      recomputed backwards.  With enough memory nothing is recomputed; with
      none it is ``"segment"``.  Float recordings unless ``tape_f64``.
    - ``"mixed"`` (F64 storage): the forward pass and its T/K checkpoints in
      FP64; backwards, each segment's checkpoint is converted to FP32 and
      recomputed by FP32 steps that record their moments (float), and the
      adjoint runs in FP32 over that recording, accumulating the inflow
      gradient in FP64.  Memory as ``"segment"``; per step the backward pass
      moves 2 x (216 + 40) bytes against the forward's 432, so fwd+bwd runs at
      about 0.45 of the forward rate.  The FP32 recomputation departs from the
      FP64 trajectory by rounding only, over at most K steps from an FP64
      state: the gradient agrees with central differences of the FP64 run to
      ~1e-7 relative (tested).

    Recordings are float64 for F64 storage in ``"segment"`` mode (the
    gradient then equals the ``"recompute"`` one) and float32 otherwise,
    unless ``tape_f64`` says.  The inflow velocity of every step is the
    lattice's inflow profile (NSE_Data_InflowProfile), set by ``forward``
    from a (3, Z, Y) float64 control tensor.  Storage F64 or F32; collision
    CUM."""

    def __init__(self, lat: D3Q27Lattice, T: int, K: int | None = None, mode: str = "segment",
                 tape_f64: bool | None = None, record_budget: int | None = None):
        if lat.cfg.pattern != _abi.PATTERN_AB:
            raise ValueError("the adjoint differentiates the A-B step: build the lattice with pattern='AB'")
        if lat.cfg.collision != _abi.COLL_CUM or lat.cfg.storage not in (_abi.STORE_F64, _abi.STORE_F32):
            raise ValueError("the adjoint supports the default cumulant with F64 or F32 storage")
        if T < 1:
            raise ValueError("T must be at least 1")
        if mode not in MODES:
            raise ValueError(f"mode must be one of {MODES}")
        f64 = lat.cfg.storage == _abi.STORE_F64
        if mode == "mixed" and not f64:
            raise ValueError("the mixed schedule runs an F64 forward pass: build the lattice with storage='f64'")
        if tape_f64 is None:
            tape_f64 = f64 and mode == "segment"
        if tape_f64 and (not f64 or mode == "mixed"):
            raise ValueError("a float64 recording needs F64 storage and an F64 recomputation")
        self.lat, self.T, self.mode, self.tape_f64 = lat, T, mode, bool(tape_f64)
        self.K = K or max(1, math.ceil(math.sqrt(T)))
        self.nseg = math.ceil(T / self.K) if mode != "tape" else 1
        el = lat.df[0].dtype
        n = lat.df[0].numel()
        dev = lat.device
        tdt = torch.float64 if self.tape_f64 else torch.float32
        # every buffer starts zeroed so that slots a step does not write hold the same values everywhere
        z = lambda: torch.zeros(n, dtype=el, device=dev)  # noqa: E731
        self.ckpt = [z() for _ in range(self.nseg)] if mode != "tape" else []
        self.pp = [z() for _ in range(2)]
        self.seg = [z() for _ in range(self.K - 1)] if mode == "recompute" else []
        ntape = {"recompute": 0, "segment": self.K, "tape": T, "hybrid": 0, "mixed": self.K}[mode]
        self.tape = torch.zeros(ntape, tape_elements(lat.d.XYZ), dtype=tdt, device=dev) if ntape else None
        if mode == "mixed":
            # the FP32 lattice of the backward pass: same geometry, map, inflow and force
            self.cfg32 = _abi.Config.make(lat.cfg.collision, _abi.STORE_F32, lat.cfg.pattern, lat.cfg.macro_real)
            self.d32 = _abi.D3Q27Data()  # refreshed from the lattice's data at each backward pass
            self.pp32 = [torch.zeros(n, dtype=torch.float32, device=dev) for _ in range(2)]
            self.lam = [torch.zeros(n, dtype=torch.float32, device=dev) for _ in range(2)]
        else:
            self.lam = [z() for _ in range(2)]
        self.generation = 0  # forward passes so far: a backward belongs to the latest one
        self.rec_first, self.rec, self.rec_offset = self.nseg, None, {}
        if mode == "hybrid":
            self._plan_recording(tdt, record_budget)
        self.profile = torch.zeros(3, lat.Z, lat.Y, dtype=torch.float64, device=dev)
        lat.set_inflow(0.0, 0.0, 0.0, profile=[self.profile[0], self.profile[1], self.profile[2]])
        for k in range(3):  # bind the control buffer itself
            lat._profiles[k] = self.profile[k]
            lat.d.inflow_profile[k] = self.profile[k].data_ptr()

    def _seg_len(self, j):
        return min(self.T, (j + 1) * self.K) - j * self.K

    def _plan_recording(self, tdt, budget):
        """Record the trailing segments that fit in `budget` bytes, keeping room
        for one segment's recording for the recomputed ones."""
        lat = self.lat
        per_step = tape_elements(lat.d.XYZ) * torch.empty((), dtype=tdt).element_size()
        if budget is None:
            torch.cuda.empty_cache()  # blocks torch caches but no longer uses count as free here
            budget = int(0.8 * torch.cuda.mem_get_info(lat.device)[0])
        steps = budget // per_step
        nrec, total = 0, 0
        while nrec < self.nseg:
            m = self._seg_len(self.nseg - 1 - nrec)
            reserve = 0 if nrec + 1 == self.nseg else self.K
            if total + m + reserve > steps:
                break
            total += m
            nrec += 1
        self.rec_first = self.nseg - nrec
        off = 0
        for j in range(self.rec_first, self.nseg):
            self.rec_offset[j] = off
            off += self._seg_len(j)
        el = tape_elements(lat.d.XYZ)
        self.rec = torch.zeros(total, el, dtype=tdt, device=lat.device) if total else None
        self.tape = torch.zeros(self.K, el, dtype=tdt, device=lat.device) if self.rec_first > 0 else None

    def memory_bytes(self) -> int:
        """Device memory the schedule holds: checkpoints, ping-pong states,
        recordings and adjoint fields."""
        bufs = list(self.ckpt) + list(self.pp) + list(self.seg) + list(self.lam) + list(getattr(self, "pp32", []))
        for t in (self.tape, self.rec):
            if t is not None:
                bufs.append(t)
        return sum(b.numel() * b.element_size() for b in bufs)

    @property
    def recorded_steps(self) -> int:
        """Steps whose moments the forward pass records (the others are recomputed)."""
        return 0 if self.rec is None else self.rec.shape[0]

    # ---- single launches ------------------------------------------------------
    def _step(self, src: torch.Tensor, dst: torch.Tensor, tape: torch.Tensor | None = None, stream=None,
              fp32: bool = False):
        lat = self.lat
        cfg, d = (self.cfg32, self.d32) if fp32 else (lat.cfg, lat.d)
        d.even_iter = 1
        d.df[0], d.df[1] = src.data_ptr(), dst.data_ptr()
        if tape is None:
            lat._check(lat.lib.tnl_lbm_d3q27_step(C.byref(cfg), C.byref(d), _abi.I64x3(0, 0, 0),
                                                  _abi.I64x3(lat.X, lat.Y, lat.Z), 0, _stream_ptr(stream)))
        else:
            lat._check(lat.lib.tnl_lbm_d3q27_step_record(C.byref(cfg), C.byref(d), _ptr(tape), int(self.tape_f64),
                                                         _stream_ptr(stream)))

    def _adjoint(self, state: torch.Tensor | None, lam_next: torch.Tensor, lam: torch.Tensor, grad: torch.Tensor,
                 tape: torch.Tensor | None = None, stream=None, fp32: bool = False):
        lat = self.lat
        cfg, d = (self.cfg32, self.d32) if fp32 else (lat.cfg, lat.d)
        if tape is None:
            d.df[0], d.df[1] = state.data_ptr(), None
            lat._check(lat.lib.tnl_lbm_d3q27_adjoint_step(C.byref(cfg), C.byref(d), _ptr(lam_next), _ptr(lam),
                                                          _ptr(grad), _stream_ptr(stream)))
        else:
            lat._check(lat.lib.tnl_lbm_d3q27_adjoint_step_recorded(C.byref(cfg), C.byref(d), _ptr(tape),
                                                                   int(self.tape_f64), _ptr(lam_next), _ptr(lam),
                                                                   _ptr(grad), _stream_ptr(stream)))

    def _refresh_fp32(self):
        """The FP32 data struct takes the lattice's geometry, map, inflow profile and force."""
        C.pointer(self.d32)[0] = self.lat.d

    def _other(self, src):
        return self.pp[1] if src is self.pp[0] else self.pp[0]

    # ---- the run ----------------------------------------------------------------
    def forward(self, inflow: torch.Tensor, f0: torch.Tensor) -> torch.Tensor:
        """Runs T steps from f0 (flat storage layout) with the given inflow
        profile; returns the buffer holding the final state (valid until the
        next call)."""
        try:
            self.generation += 1
            self.profile.copy_(inflow.reshape(3, self.lat.Z, self.lat.Y))
            if self.mode == "tape":
                src = self.pp[0]
                src.copy_(f0.reshape(-1))
                for t in range(self.T):
                    dst = self._other(src)
                    self._step(src, dst, self.tape[t])
                    src = dst
                return src
            self.ckpt[0].copy_(f0.reshape(-1))
            src = self.ckpt[0]
            for t in range(self.T):
                nxt = t + 1
                dst = self.ckpt[nxt // self.K] if nxt % self.K == 0 and nxt < self.T else self._other(src)
                j = t // self.K
                tape = self.rec[self.rec_offset[j] + t - j * self.K] if j >= self.rec_first else None
                self._step(src, dst, tape)
                src = dst
            return src
        finally:
            self.lat._bind()

    def backward(self, lam_T: torch.Tensor, grad: torch.Tensor) -> torch.Tensor:
        """lam_T = dL/d(final state); adds dL/d(inflow) to grad (3, Z, Y) float64
        and returns dL/d(initial state) (a buffer valid until the next call).
        Needs the recording of the last ``forward`` in ``"tape"`` mode."""
        try:
            cur, other = self.lam
            cur.copy_(lam_T.reshape(-1))
            if self.mode == "tape":
                for t in reversed(range(self.T)):
                    self._adjoint(None, cur, other, grad, self.tape[t])
                    cur, other = other, cur
                return cur
            for j in reversed(range(self.nseg)):
                t0 = j * self.K
                m = min(self.T, t0 + self.K) - t0
                if j >= self.rec_first:
                    for i in reversed(range(m)):
                        self._adjoint(None, cur, other, grad, self.rec[self.rec_offset[j] + i])
                        cur, other = other, cur
                elif self.mode == "mixed":
                    self._refresh_fp32()
                    src = self.pp32[0]
                    src.copy_(self.ckpt[j])  # FP64 checkpoint -> FP32 state
                    for i in range(m):
                        dst = self.pp32[1] if src is self.pp32[0] else self.pp32[0]
                        self._step(src, dst, self.tape[i], fp32=True)
                        src = dst
                    for i in reversed(range(m)):
                        self._adjoint(None, cur, other, grad, self.tape[i], fp32=True)
                        cur, other = other, cur
                elif self.mode == "recompute":
                    states = [self.ckpt[j]] + self.seg[:m - 1]
                    for i in range(m - 1):
                        self._step(states[i], states[i + 1])
                    for i in reversed(range(m)):
                        self._adjoint(states[i], cur, other, grad)
                        cur, other = other, cur
                else:
                    src = self.ckpt[j]
                    for i in range(m):
                        dst = self._other(src)
                        self._step(src, dst, self.tape[i])
                        src = dst
                    for i in reversed(range(m)):
                        self._adjoint(None, cur, other, grad, self.tape[i])
                        cur, other = other, cur
            return cur
        finally:
            self.lat._bind()


class _LatticeRun(torch.autograd.Function):
    @staticmethod
    def forward(ctx, inflow, f0, run: CheckpointedRun):
        fT = run.forward(inflow.detach(), f0.detach())
        ctx.run = run
        ctx.generation = run.generation
        ctx.shapes = (inflow.shape, f0.shape, f0.dtype)
        return fT.clone()

    @staticmethod
    def backward(ctx, gF):
        run = ctx.run
        if run.generation != ctx.generation:
            raise RuntimeError("the CheckpointedRun was run forward again before this backward pass: its checkpoints "
                               "belong to the later run (use one CheckpointedRun per graph)")
        grad = torch.zeros(3, run.lat.Z, run.lat.Y, dtype=torch.float64, device=gF.device)
        lam0 = run.backward(gF.to(run.lam[0].dtype).contiguous(), grad)
        inflow_shape, f0_shape, f0_dtype = ctx.shapes
        return grad.reshape(inflow_shape), lam0.to(f0_dtype).reshape(f0_shape), None


def lattice_run(run: CheckpointedRun, inflow: torch.Tensor, f0: torch.Tensor) -> torch.Tensor:
    """Differentiable T-step run: (inflow profile (3, Z, Y) float64, initial
    populations (flat storage layout)) -> final populations (same layout).
    Gradients flow to both inputs through the hand-written adjoint."""
    return _LatticeRun.apply(inflow, f0, run)


# ---- coupled NSE + passive scalar --------------------------------------------------

class CoupledScalarRun:
    """T coupled NSE + D3Q7 A-B steps (tnl_lbm_nse_ade_step, kernels.h:114-196)
    from given initial states, differentiable with respect to the scalar's
    source field.

    The scalar does not act on the flow and, for a given velocity, its step is
    linear in the D3Q7 populations, so the backward pass needs the velocity of
    every step and nothing of the scalar's own trajectory.  The flow is
    checkpointed every K steps (K = ceil(sqrt(T))); backwards, each segment is
    recomputed by plain D3Q27 steps that write their velocity (macro output
    into one buffer per step, the same velocity the coupled kernel hands to
    the D3Q7 update), and the D3Q7 adjoint steps (tnl_lbm_d3q7_adjoint_step)
    run over it in reverse.  Memory: T/K flow states + K velocity fields,
    packed in 3K + 1 planes: step i of a segment writes its macro output (rho,
    vx, vy, vz) three planes below step i - 1's, so that only step i - 1's
    unused density plane is overwritten.  ``host_checkpoints`` keeps the flow
    checkpoints in host memory (copied at each checkpoint and at the start of
    each segment's recomputation), for a lattice whose T/K device checkpoints
    would not fit beside the K velocity fields (FP64 at 384^3: 12.2 GB per
    state).  Any storage form; the adjoint runs in the compute type (double
    for F64, float otherwise) and accumulates the source gradient in double."""

    def __init__(self, nse: D3Q27Lattice, ade, T: int, K: int | None = None, host_checkpoints: bool = False):
        if nse.cfg.pattern != _abi.PATTERN_AB:
            raise ValueError("the adjoint differentiates the A-B step: build the lattices with pattern='AB'")
        if nse.cfg.collision != _abi.COLL_CUM:
            raise ValueError("the coupled kernel is built for the default cumulant")
        if ade.nse is not nse:
            raise ValueError("the D3Q7 lattice belongs to another NSE lattice")
        if T < 1:
            raise ValueError("T must be at least 1")
        self.nse, self.ade, self.T = nse, ade, T
        self.K = K or max(1, math.ceil(math.sqrt(T)))
        self.nseg = math.ceil(T / self.K)
        dev, n = nse.device, nse.d.XYZ
        el = nse.df[0].dtype
        mdt = torch.float64 if nse.cfg.macro_real == _abi.REAL_F64 else torch.float32
        cdt = torch.float64 if nse.cfg.storage == _abi.STORE_F64 else torch.float32
        self.host = bool(host_checkpoints)
        cdev = "cpu" if self.host else dev
        self.ckpt = [torch.zeros(27 * n, dtype=el, device=cdev) for _ in range(self.nseg)]
        self.pp = [torch.zeros(27 * n, dtype=el, device=dev) for _ in range(2)]
        self.app = [torch.zeros(7 * n, dtype=el, device=dev) for _ in range(2)]
        self.vel = torch.zeros((3 * self.K + 1) * n, dtype=mdt, device=dev)
        self.lam = [torch.zeros(7 * n, dtype=cdt, device=dev) for _ in range(2)]
        self.out = torch.zeros(7 * n, dtype=torch.float64, device=dev)
        self.source = torch.zeros(n, dtype=mdt, device=dev)
        ade.source = self.source
        ade._bind()

    def _launch(self, rc):
        self.nse._check(rc)

    def _coupled(self, fs, fd, gs, gd, stream=None):
        nse, ade = self.nse, self.ade
        nse.d.even_iter = ade.d.even_iter = 1
        nse.d.df[0], nse.d.df[1] = fs.data_ptr(), fd.data_ptr()
        ade.d.df[0], ade.d.df[1] = gs.data_ptr(), gd.data_ptr()
        self._launch(nse.lib.tnl_lbm_nse_ade_step(C.byref(nse.cfg), C.byref(nse.d), C.byref(ade.d),
                                                  _abi.I64x3(0, 0, 0), _abi.I64x3(nse.X, nse.Y, nse.Z), 0,
                                                  _stream_ptr(stream)))

    def _flow_step(self, fs, fd, ubuf, stream=None):
        nse = self.nse
        nse.d.even_iter = 1
        nse.d.df[0], nse.d.df[1] = fs.data_ptr(), fd.data_ptr()
        nse.d.macro = ubuf.data_ptr()
        self._launch(nse.lib.tnl_lbm_d3q27_step(C.byref(nse.cfg), C.byref(nse.d), _abi.I64x3(0, 0, 0),
                                                _abi.I64x3(nse.X, nse.Y, nse.Z), _abi.FLAG_OUTPUT_MACRO,
                                                _stream_ptr(stream)))

    def _macro_slot(self, i):
        """The 4 planes step i of a segment writes its macro output to."""
        n = self.nse.d.XYZ
        base = 3 * (self.K - 1 - i) * n
        return self.vel[base:base + 4 * n]

    def _adjoint(self, ubuf, lam_next, lam, grad, stream=None):
        nse = self.nse
        self._launch(nse.lib.tnl_lbm_d3q7_adjoint_step(C.byref(nse.cfg), C.byref(nse.d), C.byref(self.ade.d),
                                                       _ptr(ubuf[nse.d.XYZ:]), _ptr(lam_next), _ptr(lam),
                                                       _ptr(grad), _stream_ptr(stream)))

    def export(self, g: torch.Tensor) -> torch.Tensor:
        """A D3Q7 state (storage form) -> absolute populations, flat float64
        (a buffer valid until the next call)."""
        ade = self.ade
        saved = ade.d.df[0]
        ade.d.df[0] = g.data_ptr()
        try:
            ade._check(ade.lib.tnl_lbm_d3q7_export_f64(C.byref(ade.cfg), C.byref(ade.d), 0, _ptr(self.out),
                                                       _stream_ptr()))
        finally:
            ade.d.df[0] = saved
        return self.out

    def forward(self, source: torch.Tensor, g0: torch.Tensor, f0: torch.Tensor) -> torch.Tensor:
        """T coupled steps from (f0, g0) (flat storage layout) with the given
        source (flat XYZ in the lattice layout); returns the buffer holding
        the final D3Q7 state."""
        try:
            self.ade.refresh_rows()
            self.source.copy_(source.reshape(-1))
            self.ckpt[0].copy_(f0.reshape(-1))
            self.app[0].copy_(g0.reshape(-1))
            fs, gs = self.ckpt[0], self.app[0]
            if self.host:
                self.pp[0].copy_(self.ckpt[0])
                fs = self.pp[0]
            for t in range(self.T):
                nxt = t + 1
                at_ckpt = nxt % self.K == 0 and nxt < self.T
                other = self.pp[1] if fs is self.pp[0] else self.pp[0]
                fd = self.ckpt[nxt // self.K] if at_ckpt and not self.host else other
                gd = self.app[1] if gs is self.app[0] else self.app[0]
                self._coupled(fs, fd, gs, gd)
                if at_ckpt and self.host:
                    self.ckpt[nxt // self.K].copy_(fd)
                fs, gs = fd, gd
            return gs
        finally:
            self.nse._bind()
            self.ade._bind()

    def backward(self, lam_T: torch.Tensor, grad: torch.Tensor) -> torch.Tensor:
        """lam_T = dL/d(final absolute D3Q7 populations); adds dL/d(source) to
        grad (flat XYZ float64) and returns dL/d(initial absolute D3Q7
        populations) (a buffer valid until the next call)."""
        try:
            self.ade.refresh_rows()
            cur, other = self.lam
            cur.copy_(lam_T.reshape(-1))
            for j in reversed(range(self.nseg)):
                t0 = j * self.K
                m = min(self.T, t0 + self.K) - t0
                if self.host:
                    self.pp[0].copy_(self.ckpt[j])
                    fs = self.pp[0]
                else:
                    fs = self.ckpt[j]
                for i in range(m):
                    fd = self.pp[1] if fs is self.pp[0] else self.pp[0]
                    self._flow_step(fs, fd, self._macro_slot(i))
                    fs = fd
                for i in reversed(range(m)):
                    self._adjoint(self._macro_slot(i), cur, other, grad)
                    cur, other = other, cur
            return cur
        finally:
            self.nse._bind()
            self.ade._bind()


class _ScalarRun(torch.autograd.Function):
    @staticmethod
    def forward(ctx, source, run: CoupledScalarRun, g0, f0):
        gT = run.forward(source.detach(), g0, f0)
        ctx.run = run
        ctx.shape = source.shape
        return run.export(gT).clone()

    @staticmethod
    def backward(ctx, gG):
        run = ctx.run
        grad = torch.zeros(run.nse.d.XYZ, dtype=torch.float64, device=gG.device)
        run.backward(gG.to(run.lam[0].dtype).contiguous(), grad)
        return grad.reshape(ctx.shape), None, None, None


def scalar_run(run: CoupledScalarRun, source: torch.Tensor, g0: torch.Tensor, f0: torch.Tensor) -> torch.Tensor:
    """Differentiable coupled run: the source field (flat XYZ float64 in the
    lattice layout) -> the final absolute D3Q7 populations (flat 7*XYZ
    float64).  g0 and f0 are the initial D3Q7 and D3Q27 states in the storage
    form; the gradient flows to the source."""
    return _ScalarRun.apply(source, run, g0, f0)


def top_layer_variance(g: torch.Tensor, lat: D3Q27Lattice, layers: int) -> torch.Tensor:
    """Variance of phi = sum_q g_q over the top `layers` z-layers, in float64
    (g: flat 7*XYZ absolute populations)."""
    phi = g.view(7, lat.d.XYZ)[:, :lat.nsites].sum(0).view(lat.X, lat.Z, lat.Y)[:, lat.Z - layers:, :]
    return ((phi - phi.mean()) ** 2).mean()
