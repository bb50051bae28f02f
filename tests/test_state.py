"""The simulation driver (pytnl_lbm.state) on the CPU: the Lattice unit
helpers (lattice.h:15-156), Actions and counters (state.h:38-71), and the
control flow of execute (core.h:47-112) and State (state.hpp:1151-1590),
with a stub in place of the GPU lattice."""
import json
import math

import numpy as np
import pytest

import pytnl_lbm.state as S
from pytnl_lbm import _abi


def test_lattice_units_as_sim_1_sets_them():
    """sim_1.py: physDt = LBM_VISCOSITY / PHYS_VISCOSITY * physDl^2, so
    lbmViscosity() returns LBM_VISCOSITY; sites sit at cell centres:
    lbm2physPoint(x) = origin + (x - 1/2) physDl."""
    Y = 34
    dl = 0.41 / (Y - 2)
    nu_lbm, nu = 1e-5, 1.5e-5
    lat = S.Lattice((128, Y, Y), physOrigin=(0.1, -0.2, 0.3), physDl=dl, physDt=nu_lbm / nu * dl * dl, physViscosity=nu)
    assert math.isclose(lat.lbmViscosity(), nu_lbm, rel_tol=1e-14)
    assert math.isclose(lat.lbm2physViscosity(lat.lbmViscosity()), nu, rel_tol=1e-14)
    assert np.allclose(lat.lbm2physPoint(0.5, 0.5, 0.5), lat.physOrigin)
    assert np.allclose(lat.lbm2physPoint(3, 4, 5), lat.physOrigin + (np.array([3, 4, 5]) - 0.5) * dl)
    p = np.array([0.7, 0.01, -0.13])
    assert np.allclose(lat.lbm2physPoint(lat.phys2lbmPoint(p)), p)
    assert math.isclose(lat.phys2lbmX(lat.lbm2physX(17)), 17)
    assert math.isclose(lat.phys2lbmZ(lat.lbm2physZ(3)), 3)
    v = 1.0
    assert math.isclose(lat.lbm2physVelocity(lat.phys2lbmVelocity(v)), v)
    assert math.isclose(lat.phys2lbmVelocity(v), v * lat.physDt / dl)
    F = 2.5
    assert math.isclose(lat.lbm2physForce(lat.phys2lbmForce(F)), F)
    assert math.isclose(lat.phys2lbmForce(F), F * lat.physDt ** 2 / dl)
    assert lat.size() == (128, Y, Y) and lat.getDimensions() == (127, Y - 1, Y - 1)


def test_actions_and_counters():
    assert [a.name for a in S.Actions] == ["STAT_RESET", "STAT2_RESET", "PRINT", "OUT2D", "OUT3D", "OUT3DCUT", "PROBE1",
                                           "PROBE2", "PROBE3", "SAVESTATE", "MAX_COUNTER"]
    assert S.PRINT == 2 and S.SAVESTATE == 9
    c = S.counter()
    assert not c.action(1e9)  # period <= 0: never
    c.period = 0.5
    assert c.action(0.0)
    c.count = 3
    assert not c.action(1.49) and c.action(1.5)


class StubLattice:
    """Stands in for D3Q27Lattice: records steps, keeps a macro field."""

    def __init__(self, X, Y, Z, **kw):
        self.X, self.Y, self.Z = X, Y, Z
        self.kw = kw
        self.d = _abi.D3Q27Data.make(X, Y, Z)
        self.d.lbm_viscosity = kw.get("nu", 0.0)
        self.cfg = _abi.Config.make()
        self.iterations = 0
        self.steps = []
        self.map = None
        self.nan = False
        import torch
        self._m = [torch.ones(X, Y, Z, dtype=torch.float64)] + [torch.zeros(X, Y, Z, dtype=torch.float64)] * 3

    def set_geometry(self, g):
        self.geo = np.array(g)

    def set_equilibrium(self, *a):
        self.eq = a

    def compute_macro(self):
        pass

    def update_kernel_data(self):
        self.d.even_iter = 1 if self.iterations % 2 == 0 else 0

    def step(self, output_macro=False):
        self.update_kernel_data()
        self.steps.append((self.iterations, output_macro))
        self.iterations += 1

    def macros(self):
        return self._m

    def reduce_macro(self):
        return [1.0, 0.0, 1.0 if self.nan else 0.0, 0.0, 0.0, 1.0]

    def synchronize(self):
        pass


@pytest.fixture
def stub(monkeypatch):
    def allocate(self):
        for b in self.blocks:
            if b.lattice is None:
                b.lattice = StubLattice(*b.local, nu=self.lat.lbmViscosity())
    monkeypatch.setattr(S.LBM, "allocateDeviceData", allocate)
    monkeypatch.setattr(S.State, "estimateMemoryDemands", lambda self: True)


def lat_small():
    return S.Lattice((6, 5, 4), physDl=0.1, physDt=0.01, physViscosity=0.02)


def test_execute_runs_to_the_final_time_with_outputs(tmp_path, stub):
    class Local(S.State):
        probes = 0

        def setupBoundaries(self):
            self.nse.setBoundaryX(0, self.nse.BC.GEO_INFLOW)
            self.nse.setBoundaryX(5, self.nse.BC.GEO_OUTFLOW_EQ)
            self.nse.setMap(2, 2, 2, self.nse.BC.GEO_WALL)

        def probe1(self):
            self.probes += 1

    st = Local("t1", lat_small(), results_root=tmp_path)
    assert st.canCompute()
    st.nse.physFinalTime = 0.1  # 10 steps of physDt = 0.01, quit after step 11
    st.cnt[S.PROBE1].period = 0.05
    st.cnt[S.OUT3D].period = 0.05
    st.add2Dcut_Z(2, "midZ")
    st.cnt[S.OUT2D].period = 0.1
    S.execute(st)
    lat = st.nse.blocks[0].lattice
    assert st.nse.iterations == 11 and len(lat.steps) == 11
    assert st.flagExists("finished") and not st.canCompute()
    geo = lat.geo
    assert (geo[0] == _abi.GEO.INFLOW).all() and geo[2, 2, 2] == _abi.GEO.WALL and geo[3, 1, 1] == _abi.GEO.FLUID
    # probe and outputs at t = 0 (before the loop), 0.05, 0.10
    assert st.probes == 3
    outs = sorted(p.name for p in (st.results / "output_3D").iterdir())
    assert outs == ["output_3D_000000.vti", "output_3D_000001.vti", "output_3D_000002.vti"]
    assert len(list((st.results / "output_2D_midZ").iterdir())) == 2
    # the macro output of a step is requested exactly when an action falls due after it
    due = [it + 1 for it, om in lat.steps if om]
    assert due == [5, 10]
    from pytnl_lbm import io as lio
    f = lio.read_vti(st.results / "output_3D" / "output_3D_000001.vti")
    assert f["wall"][2, 2, 2] == _abi.GEO.WALL and f["lbm_density"].shape == (6, 5, 4)


def test_nan_terminates(tmp_path, stub):
    st = S.State("t2", lat_small(), results_root=tmp_path)
    st.nse.physFinalTime = 10.0
    st.cnt[S.PROBE2].period = 0.02
    st.SimInit()
    st.nse.blocks[0].lattice.nan = True
    st.AfterSimUpdate()
    for _ in range(5):
        st.updateKernelData()
        st.SimUpdate()
        st.AfterSimUpdate()
        if st.nse.terminate:
            break
    assert st.nse.terminate


def test_wall_time_saves_state(tmp_path, stub, monkeypatch):
    saved = []
    monkeypatch.setattr(S.State, "saveState", lambda self: (saved.append(self.nse.iterations), self.flagCreate("loadstate")))
    st = S.State("t3", lat_small(), results_root=tmp_path)
    st.nse.physFinalTime = 1e9
    st.wallTime = 1e-9
    S.execute(st)
    assert saved == [1] and st.flagExists("loadstate") and st.canCompute()


def test_second_instance_cannot_compute(tmp_path, stub):
    a = S.State("t4", lat_small(), results_root=tmp_path)
    b = S.State("t4", lat_small(), results_root=tmp_path)
    assert a.canCompute() and not b.canCompute()
