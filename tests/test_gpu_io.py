"""Checkpoint/restart and output on the GPU: a run stopped at step k,
saved (state.hpp:952-988's staging + atomic swap) and loaded into a fresh
lattice continues bit for bit like the uninterrupted run; the VTK output
holds the block's macroscopic fields."""
import numpy as np
import pytest
import torch

from pytnl_lbm import D3Q7Lattice, D3Q27Lattice, _abi, tgv_fields
from pytnl_lbm import io as lio
from pytnl_lbm._abi import ADE_GEO, GEO

pytestmark = pytest.mark.gpu


def raw(lat):
    out = []
    for t in lat.df:
        out.append((t.view(torch.int16) if t.dtype == torch.bfloat16 else t).clone())
    return out


def make(storage, pattern, geo):
    X, Y, Z = geo.shape if geo is not None else (12, 10, 8)
    lat = D3Q27Lattice(X, Y, Z, storage=storage, pattern=pattern, nu=0.02, geo=geo, store_range=0.5)
    r = tgv_fields(X, Y, Z, 0.04)
    lat.set_equilibrium_field(*r)
    if geo is not None:
        lat.set_inflow(0.02, 0.0, 0.0)
    return lat


def channel(X=12, Y=9, Z=8):
    geo = np.full((X, Y, Z), GEO.FLUID, dtype=np.int16)
    geo[:, 0] = geo[:, -1] = GEO.WALL
    geo[:, :, 0] = geo[:, :, -1] = GEO.WALL
    geo[0, 1:-1, 1:-1] = GEO.INFLOW
    geo[-1, 1:-1, 1:-1] = GEO.OUTFLOW_EQ
    return geo


@pytest.mark.parametrize("storage,pattern,geo", [("f64", "AA", None), ("f32", "AB", "channel"), ("i16dev", "AA", None),
                                                 ("bf16dev", "AA", "channel"), ("f64", "AB", None)])
def test_restart_continues_bit_for_bit(tmp_path, storage, pattern, geo):
    g = channel() if geo == "channel" else None
    ref = make(storage, pattern, g)
    run = make(storage, pattern, g)
    for _ in range(7):
        ref.step(output_macro=True)
        run.step(output_macro=True)
    torch.cuda.synchronize()
    path = lio.save_checkpoint(tmp_path / "results" / "checkpoint", [run])
    del run
    fresh = make(storage, pattern, g)  # freshly initialised: everything must come from the checkpoint
    attrs = lio.load_checkpoint(path, [fresh])
    assert attrs["LBM_iterations"] == 7 and fresh.iterations == 7
    for _ in range(6):
        ref.step(output_macro=True)
        fresh.step(output_macro=True)
    torch.cuda.synchronize()
    for a, b in zip(raw(ref), raw(fresh)):
        assert torch.equal(a, b)
    assert torch.equal(ref.macro, fresh.macro)


def test_restart_of_the_coupled_scalar(tmp_path):
    X, Y, Z = 8, 10, 6
    ageo = np.full((X, Y, Z), ADE_GEO.PERIODIC, dtype=np.int16)

    def build():
        lat = D3Q27Lattice(X, Y, Z, nu=0.02, pattern="AB")
        ade = D3Q7Lattice(lat, collision="CLBM", diffusivity=0.03, geo=ageo)
        lat.set_equilibrium_field(*tgv_fields(X, Y, Z, 0.04))
        ade.set_equilibrium(1.0)
        return lat, ade

    (l1, a1), (l2, a2) = build(), build()
    phi = torch.rand(X, Y, Z, dtype=torch.float64)
    for a in (a1, a2):
        a.set_equilibrium(phi)
    for _ in range(5):
        l1.step_coupled(a1, output_macro=True)
        l2.step_coupled(a2, output_macro=True)
    p = lio.save_checkpoint(tmp_path / "ck", [l2], ade_blocks=[a2])
    l3, a3 = build()
    lio.load_checkpoint(p, [l3], ade_blocks=[a3])
    for _ in range(4):
        l1.step_coupled(a1, output_macro=True)
        l3.step_coupled(a3, output_macro=True)
    torch.cuda.synchronize()
    assert all(torch.equal(x, y) for x, y in zip(a1.df, a3.df))
    assert torch.equal(a1.phi, a3.phi) and all(torch.equal(x, y) for x, y in zip(l1.df, l3.df))


def test_vti_output_of_a_block(tmp_path):
    lat = make("f64", "AA", None)
    for _ in range(3):
        lat.step(output_macro=True)
    torch.cuda.synchronize()
    fields = lio.macro_fields(lat)
    lio.write_vti(tmp_path / "out.vti", fields, spacing=1.0)
    back = lio.read_vti(tmp_path / "out.vti")
    rho, vx, vy, vz = (m.cpu().numpy() for m in lat.macros())
    assert np.array_equal(back["lbm_density"], rho) and np.array_equal(back["velocity_z"], vz)
    assert np.array_equal(back["lbm_density_fluctuation"], rho - 1.0)
