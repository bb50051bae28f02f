"""Pins of the oracle's adjoint of one A-B step (default cumulant), which is
computed by forward-mode automatic differentiation of the restated site map:
the dot-product identity <J v, w> = <v, J^T w> against central differences
of the oracle's own forward step, for the populations and for the inflow
velocity, on a channel with walls, inflow, both outflows and array-edge
clamps, and on a periodic box with a body force."""
import numpy as np
import pytest

import oracle as O
from pytnl_lbm import _abi
from pytnl_lbm._abi import GEO


def channel(X, Y, Z, outflow):
    geo = np.full((X, Y, Z), GEO.FLUID, dtype=np.int16)
    geo[:, 0, :] = GEO.WALL
    geo[:, -1, :] = GEO.WALL
    geo[:, :, 0] = GEO.WALL
    geo[:, :, -1] = GEO.WALL
    geo[0, 1:-1, 1:-1] = GEO.INFLOW
    geo[-1, 1:-1, 1:-1] = outflow
    return geo


def forward(lat, f, prof=None):
    """One A-B step from the state f (27, X, Y, Z); returns the new state."""
    lat.iterations = 0
    lat.df[1][:] = 0
    lat.import_f64(f, which=0)
    if prof is not None:
        lat.set_inflow(0.0, 0.0, 0.0, profile=prof)
    lat.step()
    return lat.export_f64(1)


def state(lat, seed):
    rng = np.random.default_rng(seed)
    X, Y, Z = lat.X, lat.Y, lat.Z
    r = [1 + 0.01 * rng.standard_normal((X, Y, Z))] + list(0.03 * rng.standard_normal((3, X, Y, Z)))
    lat.set_equilibrium_field(*r)
    return lat.export_f64(0) * (1 + 0.02 * rng.standard_normal((27, X, Y, Z)))


@pytest.mark.parametrize("case", ["periodic", "outflow_right", "outflow_eq"])
def test_dot_product_identity_for_the_populations(case):
    X, Y, Z = 6, 5, 4
    if case == "periodic":
        geo, force = None, (1e-4, -2e-4, 5e-5)
    else:
        geo, force = channel(X, Y, Z, GEO.OUTFLOW_RIGHT if case == "outflow_right" else GEO.OUTFLOW_EQ), (0.0, 0.0, 0.0)
    lat = O.HostLattice(X, Y, Z, pattern=_abi.PATTERN_AB, nu=0.02, geo=geo, force=force)
    lat.set_inflow(0.03, 0.01, -0.005)
    f = state(lat, 1)
    rng = np.random.default_rng(2)
    v = rng.standard_normal(f.shape) * f
    w = rng.standard_normal(f.shape)
    eps = 1e-6
    jv = (forward(lat, f + eps * v) - forward(lat, f - eps * v)) / (2 * eps)
    lat.import_f64(f, which=0)
    jtw = lat.adjoint_step(w, which=0)
    lhs, rhs = (jv * w).sum(), (v * jtw).sum()
    assert abs(lhs - rhs) < 1e-7 * abs(rhs), (lhs, rhs)


def test_dot_product_identity_for_the_inflow_velocity():
    X, Y, Z = 6, 5, 4
    lat = O.HostLattice(X, Y, Z, pattern=_abi.PATTERN_AB, nu=0.02, geo=channel(X, Y, Z, GEO.OUTFLOW_RIGHT))
    f = state(lat, 3)
    rng = np.random.default_rng(4)
    base = [0.03 + 0.01 * rng.standard_normal((Z, Y)) for _ in range(3)]
    dirs = [rng.standard_normal((Z, Y)) for _ in range(3)]
    w = rng.standard_normal(f.shape)
    eps = 1e-6
    plus = forward(lat, f, [b + eps * d for b, d in zip(base, dirs)])
    minus = forward(lat, f, [b - eps * d for b, d in zip(base, dirs)])
    lhs = (((plus - minus) / (2 * eps)) * w).sum()
    lat.set_inflow(0.0, 0.0, 0.0, profile=base)
    lat.import_f64(f, which=0)
    grad = np.zeros((3, Z, Y))
    lat.adjoint_step(w, grad=grad, which=0)
    rhs = sum((g * d).sum() for g, d in zip(grad, dirs))
    assert abs(lhs - rhs) < 1e-7 * abs(rhs), (lhs, rhs)
