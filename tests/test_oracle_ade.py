"""Pins of the oracle's D3Q7 advection-diffusion step (d3q7/*.h) and of the
coupled NSE + ADE step (kernels.h:114-196) against closed forms:

- sum of phi is conserved on a periodic box (no source);
- a Gaussian blob's variance grows by exactly 2D per step at rest, where
  D = (tau - 1/2)/4 and tau = 1/2 + 4D (d3q7/col_srt.h:17).  SRT, MRT and
  CLBM all relax the first moments with omega = 1/tau, so the rate is the
  same for all three;
- with a uniform velocity the centre of mass moves by exactly u per step,
  because the equilibrium flux is phi u (d3q7/eq.h:19-53);
- the source term adds sum S per step (w_q sums to 1);
- bounce-back walls conserve the total.
"""
import numpy as np
import pytest

import oracle as O
from pytnl_lbm import _abi
from pytnl_lbm._abi import ADE_GEO, GEO

ADE = {"SRT": _abi.ADE_COLL_SRT, "MRT": _abi.ADE_COLL_MRT, "CLBM": _abi.ADE_COLL_CLBM}
PAT = {"AA": _abi.PATTERN_AA, "AB": _abi.PATTERN_AB}


def coupled(X, Y, Z, coll, D, pattern="AA", u=(0.0, 0.0, 0.0), nu=0.01, geo=None, ade_geo=None):
    nse = O.HostLattice(X, Y, Z, pattern=PAT[pattern], nu=nu, geo=geo, threads=4)
    nse.set_equilibrium(1.0, *u)
    ade = O.HostAde(nse, collision=ADE[coll], diffusivity=D, geo=ade_geo)
    return nse, ade


def total(ade):
    return ade.export_f64().sum()


@pytest.mark.parametrize("coll", ["SRT", "MRT", "CLBM"])
@pytest.mark.parametrize("pattern", ["AA", "AB"])
def test_scalar_is_conserved_on_a_periodic_box(coll, pattern):
    X, Y, Z = 6, 7, 5
    rng = np.random.default_rng(1)
    nse = O.HostLattice(X, Y, Z, pattern=PAT[pattern], nu=0.02, threads=4)
    nse.set_equilibrium_field(1 + 0.01 * rng.standard_normal((X, Y, Z)), *(0.03 * rng.standard_normal((3, X, Y, Z))))
    ade = O.HostAde(nse, collision=ADE[coll], diffusivity=0.05)
    ade.set_equilibrium(1 + 0.2 * rng.random((X, Y, Z)))
    m0 = total(ade)
    for _ in range(7):
        nse.step_coupled(ade)
    assert abs(total(ade) / m0 - 1) < 1e-13


def moments_x(ade):
    """sum phi, mean x, variance along x of phi - 1 (profile constant in y, z)."""
    phi = ade.export_f64().sum(axis=0)[:, 0, 0] - 1.0
    x = np.arange(phi.size, dtype=np.float64)
    m0 = phi.sum()
    mean = (x * phi).sum() / m0
    return m0, mean, ((x - mean) ** 2 * phi).sum() / m0


def blob(X, Y, Z, x0, sigma, amp=0.1):
    x = np.arange(X, dtype=np.float64)[:, None, None]
    return 1.0 + amp * np.exp(-((x - x0) ** 2) / (2 * sigma**2)) * np.ones((1, Y, Z))


@pytest.mark.parametrize("coll", ["SRT", "MRT", "CLBM"])
def test_diffusion_variance_grows_by_2D_per_step(coll):
    X, Y, Z = 128, 2, 2
    D = 0.125  # tau = 1
    nse, ade = coupled(X, Y, Z, coll, D)
    ade.set_equilibrium(blob(X, Y, Z, 63.5, 4.0))
    for _ in range(20):
        nse.step_coupled(ade)
    _, _, v1 = moments_x(ade)
    for _ in range(200):
        nse.step_coupled(ade)
    _, _, v2 = moments_x(ade)
    rate = (v2 - v1) / 200
    assert abs(rate / (2 * D) - 1) < 1e-6, rate


@pytest.mark.parametrize("coll", ["SRT", "MRT", "CLBM"])
def test_advection_moves_centre_of_mass_by_u(coll):
    """The populations start at rest (f = w phi), so the flux relaxes to phi u as
    J_t = Phi u (1 - (1-omega)^t) and the centroid moves by
    u [T - (1-omega)/omega (1 - (1-omega)^T)] after T steps."""
    X, Y, Z = 128, 2, 2
    u, D, T = 0.05, 0.02, 300
    nse, ade = coupled(X, Y, Z, coll, D, u=(u, 0.0, 0.0))
    ade.set_equilibrium(blob(X, Y, Z, 40.0, 4.0))
    _, c0, _ = moments_x(ade)
    for _ in range(T):
        nse.step_coupled(ade)
    _, c1, _ = moments_x(ade)
    om = 1.0 / (0.5 + 4 * D)
    expect = u * (T - (1 - om) / om * (1 - (1 - om) ** T))
    assert abs((c1 - c0) / expect - 1) < 1e-9, (c1 - c0, expect)


def test_source_adds_its_sum_per_step():
    X, Y, Z = 6, 5, 4
    rng = np.random.default_rng(3)
    nse, ade = coupled(X, Y, Z, "SRT", 0.05, u=(0.01, 0.02, 0.0))
    ade.set_equilibrium(1.0)
    S = 1e-3 * rng.random((X, Y, Z))
    ade.set_source(S)
    m0 = total(ade)
    for _ in range(4):
        nse.step_coupled(ade)
    assert abs((total(ade) - m0) / (4 * S.sum()) - 1) < 1e-10


@pytest.mark.parametrize("pattern", ["AA", "AB"])
def test_walls_conserve_the_scalar(pattern):
    """A two-layer WALL slab (both lattices) inside a periodic box: every site
    has its true neighbours, so bounce-back must conserve the total exactly.
    (At the array edge the A-B pull clamps non-periodic neighbour indices,
    kernels.h:41-56, which is not a bijection.)"""
    X, Y, Z = 6, 5, 8
    geo = np.full((X, Y, Z), GEO.PERIODIC, dtype=np.int16)
    geo[:, :, 3:5] = GEO.WALL
    ageo = np.full((X, Y, Z), ADE_GEO.PERIODIC, dtype=np.int16)
    ageo[:, :, 3:5] = ADE_GEO.WALL
    rng = np.random.default_rng(4)
    nse, ade = coupled(X, Y, Z, "CLBM", 0.05, pattern=pattern, u=(0.02, 0.0, 0.0), geo=geo, ade_geo=ageo)
    ade.set_equilibrium(1 + 0.1 * rng.random((X, Y, Z)))
    m0 = total(ade)
    for _ in range(9):
        nse.step_coupled(ade)
    assert abs(total(ade) / m0 - 1) < 1e-13
