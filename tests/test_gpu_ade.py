"""GPU parity of the coupled NSE + D3Q7 ADE step (tnl_lbm_nse_ade_step,
replacing cudaLBMKernel<NSE,ADE>, kernels.h:114-196) against the oracle on
the same seeded inputs: D3Q27 and D3Q7 populations and the scalar output,
every step."""
import numpy as np
import pytest
import torch

import oracle as O
from lbm_common import rel_max
from pytnl_lbm import D3Q7Lattice, D3Q27Lattice, _abi
from pytnl_lbm._abi import ADE_GEO, GEO

pytestmark = pytest.mark.gpu

STORE = {"f64": _abi.STORE_F64, "f32": _abi.STORE_F32, "f32dev": _abi.STORE_F32_DEV,
         "f16dev": _abi.STORE_F16_DEV, "bf16dev": _abi.STORE_BF16_DEV, "i16dev": _abi.STORE_I16_DEV}
PAT = {"AA": _abi.PATTERN_AA, "AB": _abi.PATTERN_AB}
ADE = {"SRT": _abi.ADE_COLL_SRT, "MRT": _abi.ADE_COLL_MRT, "CLBM": _abi.ADE_COLL_CLBM}


def pair(X, Y, Z, ade_coll, storage="f64", pattern="AA", geo=None, ade_geo=None, nu=0.02, D=0.05, phi_ref=0.0,
         scale=1.0, store_range=1.0):
    mreal = _abi.REAL_F64 if storage == "f64" else _abi.REAL_F32
    g = D3Q27Lattice(X, Y, Z, storage=storage, pattern=pattern, nu=nu, geo=geo, macro_real=mreal,
                     store_range=store_range)
    h = O.HostLattice(X, Y, Z, storage=STORE[storage], pattern=PAT[pattern], nu=nu, geo=geo, macro_real=mreal,
                      store_range=store_range, threads=8)
    ga = D3Q7Lattice(g, collision=ade_coll, diffusivity=D, geo=ade_geo, phi_ref=phi_ref, store_scale=scale)
    ha = O.HostAde(h, collision=ADE[ade_coll], diffusivity=D, geo=ade_geo, phi_ref=phi_ref, store_scale=scale)
    return g, h, ga, ha


def init(g, h, ga, ha, seed, amp=0.03):
    X, Y, Z = g.X, g.Y, g.Z
    rng = np.random.default_rng(seed)
    r = [1 + 0.01 * rng.standard_normal((X, Y, Z))] + list(amp * rng.standard_normal((3, X, Y, Z)))
    g.set_equilibrium_field(*r)
    h.set_equilibrium_field(*r)
    phi = 1 + 0.2 * rng.random((X, Y, Z))
    ga.set_equilibrium(phi)
    ha.set_equilibrium(phi)
    return rng


def compare(g, h, ga, ha, tol27, tol7):
    torch.cuda.synchronize()
    assert rel_max(g.export_f64().cpu().numpy(), h.export_f64()) < tol27
    assert rel_max(ga.export_f64().cpu().numpy(), ha.export_f64()) < tol7


@pytest.mark.parametrize("ade_coll", ["SRT", "MRT", "CLBM"])
@pytest.mark.parametrize("pattern", ["AA", "AB"])
def test_coupled_periodic_f64_matches_oracle(ade_coll, pattern):
    X, Y, Z = 5, 70, 4
    g, h, ga, ha = pair(X, Y, Z, ade_coll, pattern=pattern)
    init(g, h, ga, ha, 11)
    compare(g, h, ga, ha, 1e-15, 1e-15)
    for t in range(6):
        g.step_coupled(ga, output_macro=True)
        h.step_coupled(ha, output_macro=True)
        compare(g, h, ga, ha, 1e-13, 1e-13)
        assert rel_max(ga.phi_xyz().cpu().numpy(), ha.phi_xyz()) < 1e-13, t
        for a, b in zip(g.macros(), h.macros()):
            assert rel_max(a.cpu().numpy(), b) < 1e-13


def channel(X, Y, Z):
    geo = np.full((X, Y, Z), GEO.FLUID, dtype=np.int16)
    ageo = np.full((X, Y, Z), ADE_GEO.FLUID, dtype=np.int16)
    for g, wall, nothing in ((geo, GEO.WALL, GEO.NOTHING), (ageo, ADE_GEO.WALL, ADE_GEO.NOTHING)):
        g[:, 1, :] = wall
        g[:, -2, :] = wall
        g[:, :, 1] = wall
        g[:, :, -2] = wall
        g[0] = nothing
        g[-1] = nothing
        g[:, 0] = nothing
        g[:, -1] = nothing
        g[:, :, 0] = nothing
        g[:, :, -1] = nothing
    geo[1, 2:-2, 2:-2] = GEO.INFLOW
    ageo[1, 2:-2, 2:-2] = ADE_GEO.INFLOW
    return geo, ageo


@pytest.mark.parametrize("pattern", ["AA", "AB"])
def test_coupled_channel_with_source_and_diffusivity_field(pattern):
    """Inflow (phi = 1.5), walls, outflow (OUTFLOW_EQ / OUTFLOW_RIGHT for the
    NSE, OUTFLOW_RIGHT for the A-B ADE), a per-site diffusivity and a source."""
    X, Y, Z = 12, 9, 8
    geo, ageo = channel(X, Y, Z)
    geo[-2, 2:-2, 2:-2] = GEO.OUTFLOW_EQ if pattern == "AA" else GEO.OUTFLOW_RIGHT
    if pattern == "AB":
        ageo[-2, 2:-2, 2:-2] = ADE_GEO.OUTFLOW_RIGHT
    g, h, ga, ha = pair(X, Y, Z, "CLBM", pattern=pattern, geo=geo, ade_geo=ageo)
    rng = np.random.default_rng(5)
    Dfield = 0.02 + 0.05 * rng.random((X, Y, Z))
    S = 1e-3 * rng.random((X, Y, Z))
    for lat, ad in ((g, ga), (h, ha)):
        lat.set_equilibrium(1.0, 0.02, 0.0, 0.0)
        lat.set_inflow(0.02, 0.0, 0.0)
        ad.set_equilibrium(1.0)
        ad.set_inflow(1.5)
        ad.set_diffusivity_field(Dfield)
        ad.set_source(S)
    for t in range(10):
        g.step_coupled(ga, output_macro=True)
        h.step_coupled(ha, output_macro=True)
        compare(g, h, ga, ha, 1e-13, 1e-13)
        assert rel_max(ga.phi_xyz().cpu().numpy(), ha.phi_xyz()) < 1e-13, t


@pytest.mark.parametrize("sym", [ADE_GEO.SYM_TOP, ADE_GEO.SYM_BOTTOM, ADE_GEO.SYM_LEFT, ADE_GEO.SYM_RIGHT,
                                 ADE_GEO.SYM_BACK, ADE_GEO.SYM_FRONT])
def test_ade_symmetry_codes_match_oracle(sym):
    X, Y, Z = 6, 6, 6
    ageo = np.full((X, Y, Z), ADE_GEO.PERIODIC, dtype=np.int16)
    ageo[2:4, 2:4, 2:4] = sym
    g, h, ga, ha = pair(X, Y, Z, "SRT", ade_geo=ageo)
    init(g, h, ga, ha, 6)
    for t in range(4):
        g.step_coupled(ga)
        h.step_coupled(ha)
        compare(g, h, ga, ha, 1e-13, 1e-13)


@pytest.mark.parametrize("storage,tol", [("f32", 1e-5), ("f32dev", 1e-5), ("f16dev", 3e-3), ("i16dev", 4e-4)])
def test_coupled_reduced_storage_matches_oracle(storage, tol):
    """ADE deviations stored as (f - w phi_ref) * scale in the NSE storage
    form; compared relative to the deviations' own scale."""
    X, Y, Z = 5, 70, 4
    g, h, ga, ha = pair(X, Y, Z, "MRT", storage=storage, phi_ref=1.0, scale=2.0, store_range=0.5)
    init(g, h, ga, ha, 12, amp=0.02)
    W7 = np.array([0.25] + [0.125] * 6)[:, None, None, None]
    for t in range(4):
        g.step_coupled(ga)
        h.step_coupled(ha)
        torch.cuda.synchronize()
        da = ga.export_f64().cpu().numpy() - W7
        db = ha.export_f64() - W7
        assert rel_max(da, db) < tol, (storage, t, rel_max(da, db))


@pytest.mark.parametrize("pattern", ["AA", "AB"])
def test_coupled_step_with_site_force_field(pattern):
    """computeForcing in the coupled kernel (kernels.h:162): the NSE site
    reads the per-site force; the scalar is advected by the forced flow."""
    X, Y, Z = 8, 10, 6
    g, h, ga, ha = pair(X, Y, Z, "CLBM", pattern=pattern)
    init(g, h, ga, ha, 31)
    F = 1e-4 * np.random.default_rng(32).standard_normal((3, X, Y, Z))
    g.set_force_field(*F)
    h.set_force_field(*F)
    for t in range(6):
        g.step_coupled(ga, output_macro=True)
        h.step_coupled(ha, output_macro=True)
        compare(g, h, ga, ha, 1e-13, 1e-13)
