"""Pins the oracle's lattice-level behaviour (streaming patterns, boundary
conditions, forcing) against invariants and closed forms: the A-A/A-B
equivalence, pure-streaming round trips, shear-wave viscous decay, the
Poiseuille curvature with bounce-back walls, and inflow equilibria."""
import numpy as np
import pytest

import oracle as O
from lbm_common import C, OPP, W, rel_max, stream_periodic
from pytnl_lbm import _abi
from pytnl_lbm._abi import GEO


def random_smooth(X, Y, Z, seed, amp=0.03):
    rng = np.random.default_rng(seed)
    rho = 1 + 0.01 * rng.standard_normal((X, Y, Z))
    u = amp * rng.standard_normal((3, X, Y, Z))
    return rho, u[0], u[1], u[2]


@pytest.mark.parametrize("coll", [0, 1, 2])
def test_aa_of_streamed_state_equals_ab(coll):
    """A-A initialised with S f0 reproduces the A-B pull scheme started from f0
    step by step (streaming_AA.h:79-116 vs streaming_AB.h:20-60)."""
    X, Y, Z = 5, 6, 4
    rho, ux, uy, uz = random_smooth(X, Y, Z, 11)
    ab = O.HostLattice(X, Y, Z, collision=coll, pattern=_abi.PATTERN_AB, nu=0.01)
    ab.set_equilibrium_field(rho, ux, uy, uz)
    f0 = ab.export_f64(0)
    aa = O.HostLattice(X, Y, Z, collision=coll, pattern=_abi.PATTERN_AA, nu=0.01)
    aa.import_f64(stream_periodic(f0))
    for _ in range(7):
        ab.step(output_macro=True)
        aa.step(output_macro=True)
        for a, b in zip(ab.macros(), aa.macros()):
            assert rel_max(a, b) < 1e-14


def _duct_macros(pattern, geo, T, profile):
    X, Y, Z = geo.shape
    h = O.HostLattice(X, Y, Z, pattern=pattern, nu=(0.51 - 0.5) / 3, geo=geo)
    h.set_equilibrium(1.0, 0.02, 0.0, 0.0)
    h.set_inflow(0.0, 0.0, 0.0, profile=list(profile))
    for t in range(T):
        h.step(output_macro=(t == T - 1))
    return [np.array(m) for m in h.macros()]


def test_aa_equals_ab_on_a_duct_whose_out_of_box_pulls_stay_outside():
    """On a duct with walls, an equilibrium inlet and an OUTFLOW_EQ outlet,
    A-A and A-B give the same macros from a uniform state (its own stream)
    when a wall column stands behind the outlet: the sites on the array's
    faces then send back into the box only what they pulled from inside it.
    With the outlet on the last plane itself the runs differ, because its
    pulls from beyond the box wrap in A-A (to the inlet plane) and clamp in
    A-B (kernels.h:5-57)."""
    from pytnl_lbm import channel_geometry, inflow_control
    X, Y, Z, T = 14, 8, 8, 16
    prof = inflow_control(Y, Z, 0.04, 0.005, seed=1)
    edge = channel_geometry(X, Y, Z, outflow=GEO.OUTFLOW_EQ)
    behind = edge.copy()
    behind[-1] = GEO.WALL
    behind[-2, 1:-1, 1:-1] = GEO.OUTFLOW_EQ
    for a, b in zip(_duct_macros(_abi.PATTERN_AB, behind, T, prof), _duct_macros(_abi.PATTERN_AA, behind, T, prof)):
        assert np.array_equal(a, b)
    d = max(np.abs(a - b).max() for a, b in zip(_duct_macros(_abi.PATTERN_AB, edge, T, prof),
                                                 _duct_macros(_abi.PATTERN_AA, edge, T, prof)))
    assert d > 1e-6


def test_aa_population_bookkeeping():
    """After an even step the A-A array holds post-collision values in the
    opposite slots of the same site; after the odd step they are streamed."""
    X, Y, Z = 4, 5, 3
    rho, ux, uy, uz = random_smooth(X, Y, Z, 12)
    ab = O.HostLattice(X, Y, Z, pattern=_abi.PATTERN_AB, nu=0.02)
    ab.set_equilibrium_field(rho, ux, uy, uz)
    f0 = ab.export_f64(0)
    aa = O.HostLattice(X, Y, Z, pattern=_abi.PATTERN_AA, nu=0.02)
    aa.import_f64(stream_periodic(f0))
    ab.step()
    aa.step()
    fab = ab.export_f64()       # C S f0, unstreamed
    faa = aa.export_f64()       # the same values in opposite slots
    assert rel_max(faa[OPP], fab) < 1e-15
    ab.step()
    aa.step()
    assert rel_max(aa.export_f64(), stream_periodic(ab.export_f64())) < 1e-14


def test_shear_wave_decay_rate():
    """A transverse shear wave decays as exp(-nu k^2 t) with nu the lattice
    viscosity (omega1 = 1/(3 nu + 1/2), col_cum.h:195)."""
    X, Y, Z = 2, 64, 2
    nu = 0.05
    U = 1e-4
    y = np.arange(Y)
    k = 2 * np.pi / Y
    ux = np.broadcast_to(U * np.sin(k * y)[None, :, None], (X, Y, Z)).copy()
    lat = O.HostLattice(X, Y, Z, nu=nu)
    lat.set_equilibrium_field(np.ones((X, Y, Z)), ux, 0 * ux, 0 * ux)
    amps = []
    steps = 400
    for t in range(steps):
        lat.step(output_macro=True)
        vx = lat.macros()[1][0, :, 0]
        amps.append(2 / Y * (vx * np.sin(k * y)).sum())
    amps = np.array(amps)
    t = np.arange(steps)
    rate = -np.polyfit(t[50:], np.log(amps[50:]), 1)[0]
    assert abs(rate / (nu * k * k) - 1) < 5e-3


@pytest.mark.parametrize("pattern", [_abi.PATTERN_AA, _abi.PATTERN_AB])
def test_poiseuille_curvature_with_bounce_back(pattern):
    """Force-driven channel between full-way bounce-back walls: the steady
    profile is a parabola with u'' = -F/(rho nu), independent of where the
    effective wall sits (bc.h:201-221 walls, Geier 2017 forcing)."""
    X, Y, Z = 1, 1, 26
    nu = 1.0 / 6.0
    F = 1e-6
    geo = np.full((X, Y, Z), GEO.FLUID, dtype=np.int16)
    geo[:, :, 0] = GEO.NOTHING
    geo[:, :, -1] = GEO.NOTHING
    geo[:, :, 1] = GEO.WALL
    geo[:, :, -2] = GEO.WALL
    # x and y are one site wide and periodic: mark fluid as periodic so it wraps
    geo[geo == GEO.FLUID] = GEO.PERIODIC
    lat = O.HostLattice(X, Y, Z, nu=nu, force=(F, 0.0, 0.0), geo=geo, pattern=pattern)
    lat.set_equilibrium()
    for _ in range(4000):
        lat.step()
    lat.step(output_macro=True)
    vx = lat.macros()[1][0, 0, :]
    z = np.arange(Z)
    fl = slice(3, Z - 3)
    a = np.polyfit(z[fl], vx[fl], 2)[0]
    assert abs(2 * a / (-F / nu) - 1) < 2e-3
    assert np.abs(np.polyval(np.polyfit(z[fl], vx[fl], 2), z[fl]) - vx[fl]).max() < 1e-3 * vx.max()


@pytest.mark.parametrize("pattern", [_abi.PATTERN_AA, _abi.PATTERN_AB])
def test_inflow_sites_write_equilibrium(pattern):
    """GEO_INFLOW sets rho = 1, u = inflow and f = f_eq (bc.h:86-90), no collision."""
    X, Y, Z = 6, 4, 4
    geo = np.full((X, Y, Z), GEO.PERIODIC, dtype=np.int16)
    geo[1] = GEO.INFLOW
    lat = O.HostLattice(X, Y, Z, nu=0.01, geo=geo, pattern=pattern)
    lat.set_equilibrium(1.01, 0.01, 0.0, 0.0)
    lat.set_inflow(0.03, 0.005, -0.002)
    lat.step(output_macro=True)
    f = lat.export_f64()
    feq = O.equilibrium(1.0, 0.03, 0.005, -0.002, kind=0)
    if pattern == _abi.PATTERN_AA:
        got = f[OPP][:, 1]   # even step: opposite slots of the site
    else:
        got = f[:, 1]
    assert np.abs(got - feq[:, None, None]).max() < 1e-16
    rho, vx, vy, vz = lat.macros()
    assert np.all(rho[1] == 1.0) and np.all(vx[1] == 0.03)


def test_outflow_right_is_pull_scheme_only():
    X, Y, Z = 5, 4, 4
    geo = np.full((X, Y, Z), GEO.PERIODIC, dtype=np.int16)
    geo[-1] = GEO.OUTFLOW_RIGHT
    lat = O.HostLattice(X, Y, Z, nu=0.01, geo=geo, pattern=_abi.PATTERN_AA)
    lat.set_equilibrium()
    with pytest.raises(_abi.LBMError):
        lat.step()
    lab = O.HostLattice(X, Y, Z, nu=0.01, geo=geo, pattern=_abi.PATTERN_AB)
    lab.set_equilibrium(1.0, 0.02, 0.0, 0.0)
    lab.step(output_macro=True)
    rho = lab.macros()[0]
    assert np.all(rho[-1] == 1.0)
