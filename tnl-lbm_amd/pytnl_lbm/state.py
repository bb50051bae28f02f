"""The simulation driver of the reference's Python binding, over the GPU lattice.

pytnl_lbm (pytnl_lbm/pytnl_lbm.cpp:15-73) exports the reference's C++
driver to Python: ``Lattice_3_*`` (lbm3d/lattice.h:15-156), the ``Actions``
enum and ``counter`` (state.h:38-71), ``State_*`` with its overridable hooks
(py_state.h:40-302), ``UniformDataWriter``, ``getMacroView`` and ``execute``
(core.h:47-112).  This module gives the same names and behaviour on top of
``D3Q27Lattice``, so a reference driver such as sim_NSE/sim_1.py is written
the same way: subclass ``State``, override ``setupBoundaries``,
``updateKernelVelocities``, ``outputData`` / ``getOutputDataNames`` and the
probes, set the counters' periods and call ``execute(state)``.  Python
subclassing replaces the nanobind trampoline.

One process drives one GPU, so an ``LBM`` holds one block (MPI blocks are
out of scope; the block's overlaps are zero).  Outputs go to
results_<id>/ as VTK ImageData files (io.write_vti) and checkpoints to
results_<id>/checkpoint (io.save_checkpoint), in place of ADIOS2.
"""
from __future__ import annotations

import enum
import fcntl
import logging
import os
import time
from pathlib import Path

import numpy as np

from . import _abi
from . import io as lio

log = logging.getLogger("pytnl_lbm")


# ---- Lattice (lattice.h:15-156) -------------------------------------------------------
class Lattice:
    """An equidistant 3-D lattice: its global size and the physical units.
    ``global_`` is the reference's ``global`` (a Python keyword)."""

    D = 3

    def __init__(self, global_=(0, 0, 0), physOrigin=(0.0, 0.0, 0.0), physDl=0.0, physDt=0.0, physViscosity=0.0):
        self.global_ = tuple(int(v) for v in global_)
        self.physOrigin = np.asarray(physOrigin, dtype=np.float64)
        self.physDl = float(physDl)
        self.physDt = float(physDt)
        self.physViscosity = float(physViscosity)

    # viscosity
    def lbmViscosity(self):
        return self.phys2lbmViscosity(self.physViscosity)

    def phys2lbmViscosity(self, physViscosity):
        return self.physDt / self.physDl / self.physDl * physViscosity

    def lbm2physViscosity(self, lbmViscosity):
        return self.physDl * self.physDl / self.physDt * lbmViscosity

    # coordinates: site (x, y, z) sits at the centre of its cell
    def lbm2physPoint(self, x, y=None, z=None):
        p = np.asarray(x if y is None else (x, y, z), dtype=np.float64)
        return self.physOrigin + (p - 0.5) * self.physDl

    def lbm2physX(self, x):
        return self.physOrigin[0] + (x - 0.5) * self.physDl

    def lbm2physY(self, y):
        return self.physOrigin[1] + (y - 0.5) * self.physDl

    def lbm2physZ(self, z):
        return self.physOrigin[2] + (z - 0.5) * self.physDl

    def phys2lbmPoint(self, p):
        return (np.asarray(p, dtype=np.float64) - self.physOrigin) / self.physDl + 0.5

    def phys2lbmX(self, x):
        return (x - self.physOrigin[0]) / self.physDl + 0.5

    def phys2lbmY(self, y):
        return (y - self.physOrigin[1]) / self.physDl + 0.5

    def phys2lbmZ(self, z):
        return (z - self.physOrigin[2]) / self.physDl + 0.5

    # velocity and force
    def lbm2physVelocity(self, v):
        return v / self.physDt * self.physDl

    def phys2lbmVelocity(self, v):
        return v * self.physDt / self.physDl

    def lbm2physForce(self, f):
        return f * self.physDl / self.physDt / self.physDt

    def phys2lbmForce(self, f):
        return f / self.physDl * self.physDt * self.physDt

    # mesh interface
    @staticmethod
    def getMeshDimension():
        return 3

    getDimension = getMeshDimension

    def size(self):
        return self.global_

    def getDimensions(self):
        return tuple(g - 1 for g in self.global_)

    def getOrigin(self):
        return self.physOrigin

    def getSpaceSteps(self):
        return np.full(3, self.physDl)


# ---- Actions and counters (state.h:38-71) ------------------------------------------
class Actions(enum.IntEnum):
    STAT_RESET = 0
    STAT2_RESET = 1
    PRINT = 2
    OUT2D = 3
    OUT3D = 4
    OUT3DCUT = 5
    PROBE1 = 6
    PROBE2 = 7
    PROBE3 = 8
    SAVESTATE = 9
    MAX_COUNTER = 10


globals().update({a.name: a for a in Actions})  # module-level names, as nb::enum_::export_values


class counter:
    """A periodic action: due at time t when period > 0 and t >= count * period."""

    def __init__(self, count: int = 0, period: float = -1.0):
        self.count = count
        self.period = period

    def action(self, t) -> bool:
        return self.period > 0 and t >= self.count * self.period


# ---- the LBM object and its block -------------------------------------------------
class MACRO:
    """D3Q27_MACRO_Default's quantities (d3q27/macro.h:60-67)."""
    e_rho, e_vx, e_vy, e_vz, N = 0, 1, 2, 3, 4


class BC:
    """D3Q27_BC_All::GEO codes (d3q27/bc.h:17-43) under their reference names."""


for _name in dir(_abi.GEO):
    if not _name.startswith("_"):
        setattr(BC, f"GEO_{_name}", int(getattr(_abi.GEO, _name)))


class LBMBlock:
    """One LBM_BLOCK (lbm_block.h): the host map and macro arrays beside the
    device lattice; ``data`` is the kernel data (NSE_Data_ConstInflow fields
    such as inflow_vx are set on it, as sim_1.py does)."""

    def __init__(self, nse: "LBM", bid: int):
        self.id = bid
        self.global_ = nse.lat.global_
        self.offset = (0, 0, 0)
        self.local = nse.lat.global_
        self.nse = nse
        self.lattice = None  # the D3Q27Lattice, created by LBM.allocateDeviceData
        self.hmap = np.full(self.local, _abi.GEO.FLUID, dtype=np.int16)
        self.hmacro = np.zeros((MACRO.N,) + tuple(self.local))

    @property
    def data(self):
        return self.lattice.d

    def is_distributed(self):
        return (False, False, False)

    def isLocalIndex(self, x, y, z):
        return all(0 <= v < n for v, n in zip((x, y, z), self.local))


class LBM:
    """The reference's LBM<NSE> (lbm.h): lattice, blocks, iteration count and
    physical time, geometry setters and host/device copies."""

    BC = BC
    MACRO = MACRO

    def __init__(self, lat: Lattice, *, collision="CUM", storage="f64", pattern="AA", macro_real=None,
                 device="cuda:0", force=(0.0, 0.0, 0.0)):
        self.lat = lat
        self.iterations = 0
        self.startIterations = 0
        self.physFinalTime = 1e30
        self.physCharLength = 0.0
        self.physStartTime = 0.0
        self.terminate = False
        self.rank, self.nproc, self.total_blocks = 0, 1, 1
        self._kind = dict(collision=collision, storage=storage, pattern=pattern, macro_real=macro_real,
                          device=device, force=force)
        self.blocks = [LBMBlock(self, 0)]

    def physTime(self):
        return self.lat.physDt * self.iterations

    def Re(self, physvel):
        return abs(physvel) * self.physCharLength / self.lat.physViscosity

    # geometry on the host map (lbm.hpp:76-116)
    def setMap(self, x, y, z, value):
        for b in self.blocks:
            if b.isLocalIndex(x, y, z):
                b.hmap[x, y, z] = value

    def setBoundaryX(self, x, value):
        for b in self.blocks:
            if 0 <= x < b.local[0]:
                b.hmap[x, :, :] = value

    def setBoundaryY(self, y, value):
        for b in self.blocks:
            if 0 <= y < b.local[1]:
                b.hmap[:, y, :] = value

    def setBoundaryZ(self, z, value):
        for b in self.blocks:
            if 0 <= z < b.local[2]:
                b.hmap[:, :, z] = value

    def resetMap(self, geo_type):
        for b in self.blocks:
            b.hmap[...] = geo_type

    # device data
    def allocateDeviceData(self):
        from .lattice import D3Q27Lattice

        k = self._kind
        for b in self.blocks:
            if b.lattice is None:
                X, Y, Z = b.local
                b.lattice = D3Q27Lattice(X, Y, Z, collision=k["collision"], storage=k["storage"], pattern=k["pattern"],
                                         macro_real=k["macro_real"], nu=self.lat.lbmViscosity(), force=k["force"],
                                         device=k["device"])

    def allocateHostData(self):
        pass  # the host arrays live with the blocks from the start

    def copyMapToDevice(self):
        for b in self.blocks:
            b.lattice.set_geometry(b.hmap)

    def copyMapToHost(self):
        for b in self.blocks:
            if b.lattice.map is not None:
                X, Y, Z = b.local
                b.hmap[...] = b.lattice.map.view(X, Z, Y).permute(0, 2, 1).cpu().numpy()

    def copyMacroToHost(self):
        for b in self.blocks:
            for k, m in enumerate(b.lattice.macros()):
                b.hmacro[k] = m.cpu().numpy()

    def copyMacroToDevice(self):
        import torch

        for b in self.blocks:
            for k, m in enumerate(b.lattice.macros()):
                m.copy_(torch.as_tensor(b.hmacro[k], dtype=m.dtype, device=m.device))

    def setEquilibrium(self, rho, vx, vy, vz):
        for b in self.blocks:
            b.lattice.set_equilibrium(rho, vx, vy, vz)

    def computeInitialMacro(self):
        for b in self.blocks:
            b.lattice.compute_macro()


def getMacroView(hmacro, quantity: int):
    """getMacroView (defs.h:179-240) over a block's host macro array: the
    [x, y, z] array of one quantity, without a copy."""
    return hmacro[int(quantity)]


class UniformDataWriter:
    """Collects the fields of one output (UniformDataWriter.hpp:20-110) and
    writes them as one VTK ImageData file on close."""

    def __init__(self, path, lat: Lattice, begin, end):
        self.path = Path(path)
        self.lat = lat
        self.begin, self.end = tuple(begin), tuple(end)
        self.fields = {}
        self.time = None

    def write(self, name, src, begin=None, end=None):
        if not isinstance(name, str):
            raise TypeError("field name must be a string")
        if np.isscalar(src) and begin is None:
            self.time = float(src) if name == "TIME" else self.time
            return
        b = tuple(begin) if begin is not None else self.begin
        e = tuple(end) if end is not None else self.end
        if callable(src):
            vals = np.empty(tuple(e[i] - b[i] for i in range(3)))
            for x in range(b[0], e[0]):
                for y in range(b[1], e[1]):
                    for z in range(b[2], e[2]):
                        vals[x - b[0], y - b[1], z - b[2]] = src(x, y, z)
        else:
            vals = np.asarray(src)[b[0]:e[0], b[1]:e[1], b[2]:e[2]]
        self.fields[name] = vals

    def close(self):
        self.path.parent.mkdir(parents=True, exist_ok=True)
        origin = self.lat.lbm2physPoint(*self.begin) - 0.5 * self.lat.physDl if self.lat.physDl else np.zeros(3)
        lio.write_vti(self.path, self.fields, spacing=self.lat.physDl or 1.0, origin=tuple(origin), time=self.time)


# ---- State (state.h, state.hpp) ----------------------------------------------------
class State:
    """The simulation state of one run (State<NSE>): the LBM object, the
    counters ``cnt`` indexed by ``Actions``, output cuts, flags and
    checkpoints under results_<id>/.  Override the hooks in a subclass."""

    def __init__(self, id: str, lat: Lattice, *, collision="CUM", storage="f64", pattern="AA", macro_real=None,
                 device="cuda:0", force=(0.0, 0.0, 0.0), results_root="."):
        self.id = id
        self.nse = LBM(lat, collision=collision, storage=storage, pattern=pattern, macro_real=macro_real,
                       device=device, force=force)
        self.cnt = [counter() for _ in range(Actions.MAX_COUNTER)]
        self.wallTime = -1.0
        self.results = Path(results_root) / f"results_{id}"
        self.results.mkdir(parents=True, exist_ok=True)
        self.cuts2D = []  # (type, name, position, cycle): probe2Dcut, state.h:29-36
        self.cuts3D = []  # (ox, oy, oz, lx, ly, lz, name, cycle): probe3Dcut, state.h:20-27
        self._t0 = time.perf_counter()
        self._glups_prev = (0.0, 0)
        self.timers = {k: 0.0 for k in ("SimInit", "SimUpdate", "AfterSimUpdate", "compute")}
        self._lock = open(self.results / "lock", "w")
        try:
            fcntl.flock(self._lock, fcntl.LOCK_EX | fcntl.LOCK_NB)
            self._locked = True
        except OSError:
            self._locked = False

    @property
    def counters(self):
        return self.cnt

    @property
    def physCharLength(self):
        return self.nse.physCharLength

    @physCharLength.setter
    def physCharLength(self, v):
        self.nse.physCharLength = v

    @property
    def physFinalTime(self):
        return self.nse.physFinalTime

    @physFinalTime.setter
    def physFinalTime(self, v):
        self.nse.physFinalTime = v

    @property
    def total_blocks(self):
        return self.nse.total_blocks

    # ---- flags (state.hpp:149-211) ---------------------------------------------------
    def flagCreate(self, name):
        (self.results / f"flag.{name}").touch()

    def flagDelete(self, name):
        p = self.results / f"flag.{name}"
        if p.exists():
            p.unlink()

    def flagExists(self, name):
        return (self.results / f"flag.{name}").exists()

    def canCompute(self):
        if not self._locked:
            log.warning("failed to lock results_%s: is another instance running?", self.id)
            return False
        if self.flagExists("loadstate"):
            return True
        if self.flagExists("finished"):
            log.info("results_%s is in the finished state, nothing to compute", self.id)
            return False
        if self.flagExists("terminated"):
            log.warning("results_%s is in the terminated state, nothing to compute", self.id)
            return False
        return True

    # ---- timing ---------------------------------------------------------------------
    def getWallTime(self, collective=False):
        return time.perf_counter() - self._t0

    def wallTimeReached(self):
        return self.wallTime > 0 and self.getWallTime() >= self.wallTime

    # ---- cuts -----------------------------------------------------------------------
    def add2Dcut_X(self, x, name):
        self.cuts2D.append([0, name, int(x), 0])

    def add2Dcut_Y(self, y, name):
        self.cuts2D.append([1, name, int(y), 0])

    def add2Dcut_Z(self, z, name):
        self.cuts2D.append([2, name, int(z), 0])

    def add3Dcut(self, ox, oy, oz, lx, ly, lz, name):
        self.cuts3D.append([int(ox), int(oy), int(oz), int(lx), int(ly), int(lz), name, 0])

    # ---- hooks with the reference's defaults -------------------------------------------
    def estimateMemoryDemands(self):
        """Enough free device memory for the populations, macro and map arrays."""
        import torch

        k = self.nse._kind
        el = {"f64": 8, "f32": 4, "f32dev": 4}.get(k["storage"], 2)
        copies = 2 if k["pattern"] == "AB" else 1
        need = sum(int(np.prod(b.local)) for b in self.nse.blocks) * (27 * el * copies + 4 * 8 + 2)
        free, _ = torch.cuda.mem_get_info(torch.device(k["device"]))
        return need < free

    def resetDFs(self):
        self.nse.setEquilibrium(1.0, 0.0, 0.0, 0.0)

    def setupBoundaries(self):
        pass

    def reset(self):
        self.resetDFs()
        self.nse.resetMap(BC.GEO_FLUID)
        self.setupBoundaries()
        self.nse.copyMapToDevice()
        self.nse.computeInitialMacro()
        self.nse.copyMacroToHost()

    def SimInit(self):
        t = time.perf_counter()
        if not self.estimateMemoryDemands():
            raise RuntimeError("not enough device memory for the lattice")
        self.nse.allocateHostData()
        for c in range(Actions.MAX_COUNTER):
            self.cnt[c].count = 0
        self.cnt[Actions.SAVESTATE].count = 1  # skip the initial save
        self.nse.iterations = 0
        self.nse.allocateDeviceData()
        if self.flagExists("loadstate"):
            self.loadState()
        else:
            self.reset()
        self._glups_prev = (self.getWallTime(), self.nse.iterations)
        self.timers["SimInit"] += time.perf_counter() - t

    def updateKernelData(self):
        """even_iter and the A-B rotation follow the iteration count, the
        viscosity follows the lattice units (state.hpp:1627-1634, lbm.hpp:307-323)."""
        for b in self.nse.blocks:
            b.lattice.iterations = self.nse.iterations
            b.lattice.d.lbm_viscosity = self.nse.lat.lbmViscosity()
            b.lattice.update_kernel_data()

    def updateKernelVelocities(self):
        pass

    def computeBeforeLBMKernel(self):
        pass

    def computeAfterLBMKernel(self):
        pass

    def probe1(self):
        pass

    def probe2(self):
        pass

    def probe3(self):
        pass

    def statReset(self):
        pass

    def stat2Reset(self):
        pass

    def getOutputDataNames(self):
        return ["lbm_density", "velocity_x", "velocity_y", "velocity_z"]

    def outputData(self, writer, block, begin, end):
        writer.write("lbm_density", getMacroView(block.hmacro, MACRO.e_rho), begin, end)
        lat = self.nse.lat
        for name, q in (("velocity_x", MACRO.e_vx), ("velocity_y", MACRO.e_vy), ("velocity_z", MACRO.e_vz)):
            v = getMacroView(block.hmacro, q)
            writer.write(name, lat.lbm2physVelocity(v) if lat.physDt else v, begin, end)

    def copyAllToDevice(self):
        self.nse.copyMapToDevice()
        self.nse.copyMacroToDevice()

    def copyAllToHost(self):
        self.nse.copyMapToHost()
        self.nse.copyMacroToHost()

    def checkpointStateLocal(self, mode):
        pass

    # ---- the step (state.hpp:1235-1409) -------------------------------------------------
    def SimUpdate(self):
        t = time.perf_counter()
        for b in self.nse.blocks:
            if b.lattice.d.lbm_viscosity == 0:
                log.error("LBM viscosity is 0")
                self.nse.terminate = True
                return
        self.computeBeforeLBMKernel()
        self.nse.iterations += 1
        now = self.nse.physTime()
        compute_macro = any(self.cnt[c].action(now) for c in range(Actions.MAX_COUNTER)
                            if c not in (Actions.PRINT, Actions.SAVESTATE))
        tc = time.perf_counter()
        for b in self.nse.blocks:
            b.lattice.step(output_macro=compute_macro)
        for b in self.nse.blocks:
            b.lattice.synchronize()
        self.timers["compute"] += time.perf_counter() - tc
        self.timers["SimUpdate"] += time.perf_counter() - t

    # ---- after the step (state.hpp:1419-1590) --------------------------------------------
    def AfterSimUpdate(self):
        t = time.perf_counter()
        now = self.nse.physTime()
        act = {c: self.cnt[c].action(now) for c in range(Actions.MAX_COUNTER)}
        copy_macro = any(act[c] for c in act if c not in (Actions.PRINT, Actions.SAVESTATE, Actions.PROBE1))
        if copy_macro:
            self.nse.copyMacroToHost()
        self.computeAfterLBMKernel()
        write_info = any(act[c] for c in (Actions.PRINT, Actions.OUT2D, Actions.OUT3D, Actions.OUT3DCUT,
                                          Actions.PROBE1, Actions.PROBE2, Actions.PROBE3))
        if write_info:
            self.cnt[Actions.PRINT].count += 1
        nan_detected = False
        if self.nse.iterations > 1 and copy_macro:
            for b in self.nse.blocks:
                if b.lattice.reduce_macro()[2] > 0:
                    log.error("NaN detected in block %d", b.id)
                    nan_detected = True
            if nan_detected:
                log.error("detected NaN, terminating the simulation")
                self.nse.terminate = True
                self.nse.copyMacroToHost()
        for c, fn in ((Actions.PROBE1, self.probe1), (Actions.PROBE2, self.probe2), (Actions.PROBE3, self.probe3)):
            if act[c]:
                fn()
                self.cnt[c].count += 1
        if act[Actions.OUT3DCUT]:
            self.write3Dcut(now, self.cnt[Actions.OUT3DCUT].count)
            self.cnt[Actions.OUT3DCUT].count += 1
        if act[Actions.OUT2D] or nan_detected:
            self.write2D(now, self.cnt[Actions.OUT2D].count)
            self.cnt[Actions.OUT2D].count += 1
        if act[Actions.OUT3D] or nan_detected:
            self.write3D(now, self.cnt[Actions.OUT3D].count)
            self.cnt[Actions.OUT3D].count += 1
        if act[Actions.STAT_RESET]:
            self.statReset()
            self.nse.copyMacroToDevice()
            self.cnt[Actions.STAT_RESET].count += 1
        if act[Actions.STAT2_RESET]:
            self.stat2Reset()
            self.nse.copyMacroToDevice()
            self.cnt[Actions.STAT2_RESET].count += 1
        if write_info and self.nse.iterations > 1:
            wall = self.getWallTime()
            t_prev, it_prev = self._glups_prev
            lups = (self.nse.iterations - it_prev) / max(1e-6, wall - t_prev) * float(np.prod(self.nse.lat.global_))
            self._glups_prev = (wall, self.nse.iterations)
            log.info("GLUPS=%.3f iter=%d t=%.3fs dt=%.2e lbmVisc=%.2e WT=%.0fs", lups * 1e-9, self.nse.iterations,
                     now, self.nse.lat.physDt, self.nse.lat.lbmViscosity(), wall)
        self.timers["AfterSimUpdate"] += time.perf_counter() - t

    def AfterSimFinished(self):
        it = self.nse.iterations - self.nse.startIterations
        if it > 1:
            log.info("total walltime %.1f s, SimInit %.1f s, SimUpdate %.1f s, AfterSimUpdate %.1f s, compute %.1f s",
                     self.getWallTime(), self.timers["SimInit"], self.timers["SimUpdate"],
                     self.timers["AfterSimUpdate"], self.timers["compute"])

    def waitForPendingIO(self):
        pass  # outputs are written synchronously

    # ---- outputs --------------------------------------------------------------------
    def write3D(self, t, cycle):
        b = self.nse.blocks[0]
        w = UniformDataWriter(self.results / "output_3D" / f"output_3D_{cycle:06d}.vti", self.nse.lat, (0, 0, 0), b.local)
        w.write("TIME", t)
        w.write("wall", b.hmap, (0, 0, 0), b.local)
        self.outputData(w, b, (0, 0, 0), b.local)
        w.close()

    def write3Dcut(self, t, cycle):
        b = self.nse.blocks[0]
        for cut in self.cuts3D:
            ox, oy, oz, lx, ly, lz, name, _ = cut
            begin, end = (ox, oy, oz), (ox + lx, oy + ly, oz + lz)
            w = UniformDataWriter(self.results / f"output_3Dcut_{name}" / f"{cycle:06d}.vti", self.nse.lat, begin, end)
            w.write("TIME", t)
            w.write("wall", b.hmap, begin, end)
            self.outputData(w, b, begin, end)
            w.close()
            cut[7] += 1

    def write2D(self, t, cycle):
        b = self.nse.blocks[0]
        for cut in self.cuts2D:
            typ, name, pos, _ = cut
            begin, end = [0, 0, 0], list(b.local)
            begin[typ], end[typ] = pos, pos + 1
            w = UniformDataWriter(self.results / f"output_2D_{name}" / f"{cycle:06d}.vti", self.nse.lat, begin, end)
            w.write("TIME", t)
            w.write("wall", b.hmap, begin, end)
            self.outputData(w, b, begin, end)
            w.close()
            cut[3] += 1

    # ---- checkpoints (state.hpp:891-1010) ------------------------------------------------
    def saveState(self):
        lio.save_checkpoint(self.results / "checkpoint", [b.lattice for b in self.nse.blocks], state=self)
        self.checkpointStateLocal("write")
        self.flagCreate("loadstate")

    def loadState(self):
        attrs = lio.load_checkpoint(self.results / "checkpoint", [b.lattice for b in self.nse.blocks], state=self)
        self.nse.iterations = int(attrs["LBM_iterations"])
        self.nse.startIterations = self.nse.iterations
        self.nse.physStartTime = self.nse.physTime()
        self.nse.copyMapToHost()
        self.nse.copyMacroToHost()
        self.checkpointStateLocal("read")


def execute(state: State):
    """The reference's main loop (core.h:47-112) on one process."""
    state.SimInit()
    state.AfterSimUpdate()
    quit_ = False
    while not quit_:
        state.updateKernelData()
        state.updateKernelVelocities()
        state.SimUpdate()
        state.AfterSimUpdate()
        if state.wallTimeReached():
            state.copyAllToHost()
            log.info("maximum wall time reached")
            state.saveState()
            quit_ = True
        elif state.cnt[Actions.SAVESTATE].action(state.getWallTime(True)):
            state.copyAllToHost()
            state.saveState()
            state.cnt[Actions.SAVESTATE].count += 1
        if state.nse.physTime() > state.nse.physFinalTime:
            log.info("physFinalTime reached")
            quit_ = True
            state.flagCreate("finished")
            state.flagDelete("loadstate")
        if state.nse.terminate:
            log.info("terminate flag triggered")
            quit_ = True
            state.flagCreate("terminated")
            state.flagDelete("loadstate")
    state.AfterSimFinished()
