"""GPU parity of the alternative D3Q27 operators -- MRT_LES (col_mrt.h) and
KBC N1-N4 / C1-C4 (col_kbc_n.h, col_kbc_c.h) -- against the oracle on the
same seeded inputs, every step: periodic boxes with a body force, the
map-reading kernel in both streaming patterns, and channel boundaries whose
equilibria are the operators' own (EQ for MRT/N1/C1, EQ_ENTROPIC else)."""
import numpy as np
import pytest
import torch

import oracle as O
from lbm_common import rel_max
from pytnl_lbm import D3Q27Lattice, _abi
from pytnl_lbm._abi import GEO

pytestmark = pytest.mark.gpu

ALT = {"MRT": _abi.COLL_MRT_LES, "N1": _abi.COLL_KBC_N1, "N2": _abi.COLL_KBC_N2, "N3": _abi.COLL_KBC_N3,
       "N4": _abi.COLL_KBC_N4, "C1": _abi.COLL_KBC_C1, "C2": _abi.COLL_KBC_C2, "C3": _abi.COLL_KBC_C3,
       "C4": _abi.COLL_KBC_C4}
STORE = {"f64": _abi.STORE_F64, "f32": _abi.STORE_F32}
PAT = {"AA": _abi.PATTERN_AA, "AB": _abi.PATTERN_AB}


def fields(X, Y, Z, seed, amp=0.03):
    rng = np.random.default_rng(seed)
    rho = 1 + 0.01 * rng.standard_normal((X, Y, Z))
    u = amp * rng.standard_normal((3, X, Y, Z))
    return rho, u[0], u[1], u[2]


def pair(X, Y, Z, name, storage="f64", pattern="AA", geo=None, nu=0.01, force=(0.0, 0.0, 0.0)):
    mreal = _abi.REAL_F64 if storage == "f64" else _abi.REAL_F32
    g = D3Q27Lattice(X, Y, Z, collision=ALT[name], storage=storage, pattern=pattern, nu=nu, force=force, geo=geo,
                     macro_real=mreal)
    h = O.HostLattice(X, Y, Z, collision=ALT[name], storage=STORE[storage], pattern=PAT[pattern], nu=nu, force=force,
                      geo=geo, macro_real=mreal, threads=8)
    return g, h


def perturb(g, h, seed, eps=0.02):
    """Equilibrium of the random fields plus a seeded non-equilibrium part (the
    KBC stabiliser is 0/0 at an exact equilibrium)."""
    r = fields(g.X, g.Y, g.Z, seed)
    h.set_equilibrium_field(*r)
    f = h.export_f64()
    f = f * (1 + eps * np.random.default_rng(seed + 100).standard_normal(f.shape))
    h.import_f64(f)
    g.import_f64(f)


def cmp_pops(g, h):
    return rel_max(g.export_f64().cpu().numpy(), h.export_f64())


@pytest.mark.parametrize("name", list(ALT))
@pytest.mark.parametrize("shape", [(6, 7, 5), (3, 70, 4)])
def test_periodic_f64_matches_oracle(name, shape):
    X, Y, Z = shape
    g, h = pair(X, Y, Z, name, force=(1e-5, -2e-5, 3e-6))
    perturb(g, h, 1)
    for t in range(6):
        g.step(output_macro=True)
        h.step(output_macro=True)
        torch.cuda.synchronize()
        assert cmp_pops(g, h) < 1e-13, (name, shape, t, cmp_pops(g, h))
        for a, b in zip(g.macros(), h.macros()):
            assert rel_max(a.cpu().numpy(), b) < 1e-13


@pytest.mark.parametrize("name", ["MRT", "N3", "C2"])
@pytest.mark.parametrize("pattern", ["AA", "AB"])
def test_general_path_matches(name, pattern):
    X, Y, Z = 5, 9, 4
    geo = np.full((X, Y, Z), GEO.PERIODIC, dtype=np.int16)
    g, h = pair(X, Y, Z, name, pattern=pattern, geo=geo)
    perturb(g, h, 2)
    for t in range(6):
        g.step(output_macro=True)
        h.step(output_macro=True)
        assert cmp_pops(g, h) < 1e-13, (name, pattern, t)


def channel_geo(X, Y, Z, outflow):
    geo = np.full((X, Y, Z), GEO.FLUID, dtype=np.int16)
    geo[:, 1, :] = GEO.WALL
    geo[:, -2, :] = GEO.WALL
    geo[:, :, 1] = GEO.WALL
    geo[:, :, -2] = GEO.WALL
    geo[1, 2:-2, 2:-2] = GEO.INFLOW
    geo[-2, 2:-2, 2:-2] = outflow
    for sl in (np.s_[0], np.s_[-1], np.s_[:, 0], np.s_[:, -1], np.s_[:, :, 0], np.s_[:, :, -1]):
        geo[sl] = GEO.NOTHING
    return geo


@pytest.mark.parametrize("name,pattern", [("MRT", "AA"), ("N1", "AB"), ("N2", "AA"), ("C4", "AB")])
def test_channel_bcs_match_oracle(name, pattern):
    X, Y, Z = 12, 9, 8
    geo = channel_geo(X, Y, Z, GEO.OUTFLOW_EQ if pattern == "AA" else GEO.OUTFLOW_RIGHT)
    g, h = pair(X, Y, Z, name, pattern=pattern, geo=geo, nu=0.02)
    for lat in (g, h):
        lat.set_equilibrium(1.0, 0.02, 0.0, 0.0)
        lat.set_inflow(0.02, 0.001, 0.0)
    # the walls and ghosts keep their equilibrium; perturb so that the stabiliser is defined
    f = h.export_f64() * (1 + 0.01 * np.random.default_rng(3).standard_normal((27, X, Y, Z)))
    h.import_f64(f)
    g.import_f64(f)
    for t in range(8):
        g.step(output_macro=True)
        h.step(output_macro=True)
        assert cmp_pops(g, h) < 1e-12, (name, t, cmp_pops(g, h))


@pytest.mark.parametrize("name", ["MRT", "N1", "N4", "C1", "C4"])
def test_f32_matches_oracle(name):
    X, Y, Z = 6, 70, 5
    g, h = pair(X, Y, Z, name, storage="f32")
    perturb(g, h, 4)
    for t in range(4):
        g.step(output_macro=True)
        h.step(output_macro=True)
        assert cmp_pops(g, h) < 2e-5, (name, t, cmp_pops(g, h))


@pytest.mark.parametrize("name,kind", [("MRT", 1), ("N1", 1), ("N2", 2), ("C3", 2)])
def test_set_equilibrium_uses_the_operators_equilibrium(name, kind):
    g = D3Q27Lattice(3, 4, 5, collision=ALT[name], nu=0.01)
    g.set_equilibrium(1.02, 0.05, -0.04, 0.03)
    np.testing.assert_allclose(g.export_f64()[:, 1, 2, 3].cpu().numpy(),
                               O.equilibrium(1.02, 0.05, -0.04, 0.03, kind=kind), rtol=1e-14)
