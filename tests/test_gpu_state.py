"""The reference-style driver on the GPU: a State subclass set up as
sim_NSE/sim_1.py sets up its channel (setupBoundaries, updateKernelVelocities,
counters, execute) steps exactly the lattice a hand-driven D3Q27Lattice
steps, bit for bit; a run stopped by its wall-time limit saves a checkpoint,
and a new State with the same id resumes it (SimInit's loadstate path) to the
same final populations as an uninterrupted run."""
import gc

import numpy as np
import pytest
import torch

from pytnl_lbm import D3Q27Lattice, GEO
from pytnl_lbm import state as S

pytestmark = pytest.mark.gpu

X, Y, Z = 24, 12, 10
U_IN = 0.03


class Channel(S.State):
    def setupBoundaries(self):
        nse = self.nse
        gx, gy, gz = nse.lat.global_
        nse.setBoundaryX(0, nse.BC.GEO_INFLOW)
        nse.setBoundaryX(gx - 1, nse.BC.GEO_OUTFLOW_EQ)
        for y in (0, gy - 1):
            nse.setBoundaryY(y, nse.BC.GEO_WALL)
        for z in (0, gz - 1):
            nse.setBoundaryZ(z, nse.BC.GEO_WALL)
        nse.setMap(8, 5, 5, nse.BC.GEO_WALL)

    def updateKernelVelocities(self):
        for block in self.nse.blocks:
            block.data.inflow_vx = U_IN
            block.data.inflow_vy = 0.0
            block.data.inflow_vz = 0.0


def lattice_units():
    return S.Lattice((X, Y, Z), physDl=0.01, physDt=0.001, physViscosity=0.01 * 0.01 * 0.01 / 0.001 * 10)


def hand_driven(steps, pattern):
    geo = np.full((X, Y, Z), GEO.FLUID, dtype=np.int16)
    geo[0] = GEO.INFLOW
    geo[-1] = GEO.OUTFLOW_EQ
    geo[:, 0] = geo[:, -1] = GEO.WALL
    geo[:, :, 0] = geo[:, :, -1] = GEO.WALL
    geo[8, 5, 5] = GEO.WALL
    lat = D3Q27Lattice(X, Y, Z, pattern=pattern, nu=lattice_units().lbmViscosity(), geo=geo)
    lat.set_equilibrium(1.0, 0.0, 0.0, 0.0)
    lat.set_inflow(U_IN, 0.0, 0.0)
    for _ in range(steps):
        lat.step()
    torch.cuda.synchronize()
    return lat


@pytest.mark.parametrize("pattern", ["AA", "AB"])
def test_execute_steps_the_same_lattice(tmp_path, pattern):
    st = Channel("chan", lattice_units(), pattern=pattern, results_root=tmp_path)
    assert st.canCompute()
    st.nse.physFinalTime = 0.0205  # physTime > final after 21 steps
    st.cnt[S.OUT3D].period = 0.01
    st.cnt[S.PROBE1].period = 0.005
    S.execute(st)
    assert st.nse.iterations == 21 and st.flagExists("finished")
    ref = hand_driven(21, pattern)
    got = st.nse.blocks[0].lattice
    assert all(torch.equal(a, b) for a, b in zip(ref.df, got.df))
    assert len(list((st.results / "output_3D").iterdir())) == 3
    rho = st.nse.blocks[0].hmacro[0]
    assert np.isfinite(rho).all() and abs(rho[1:-1, 1:-1, 1:-1].mean() - 1) < 0.05


def test_wall_time_checkpoint_and_resume(tmp_path):
    st = Channel("resume", lattice_units(), results_root=tmp_path)
    st.nse.physFinalTime = 0.0305
    st.wallTime = 1e-9  # stop after the first step: save and quit
    S.execute(st)
    stopped_at = st.nse.iterations
    assert st.flagExists("loadstate") and (st.results / "checkpoint" / "attributes.json").exists()
    del st
    gc.collect()
    st2 = Channel("resume", lattice_units(), results_root=tmp_path)
    assert st2.canCompute()
    S.execute(st2)
    assert st2.nse.iterations == 31 and stopped_at < 31
    ref = hand_driven(31, "AA")
    assert all(torch.equal(a, b) for a, b in zip(ref.df, st2.nse.blocks[0].lattice.df))
