"""Pins the oracle's per-site force (MACRO::computeForcing, d3q27/macro.h:50-52,
called before the collision at kernels.h:103): the field is added to the
homogeneous force of copyQuantities (macro.h:80-86), and every collision
operator adds the site's force to its momentum (the cumulants through
u = (j + F/2)/rho and ks_100 = -k_100, col_cum.h:371-373; KBC and MRT through
their forcing terms).  So after one A-B step from a random state, the momentum
at each site minus the momentum of the populations it pulled is exactly that
site's force, which pins the field's indexing (a transposed or shifted plane
fails); and a constant field equals the same homogeneous force bit for bit."""
import numpy as np
import pytest

import oracle as O
from lbm_common import C, CX, CY, CZ
from pytnl_lbm import _abi

X, Y, Z = 7, 6, 5
NU = 0.02


def lattice(coll, storage=_abi.STORE_F64, force=(0.0, 0.0, 0.0), seed=1):
    mreal = _abi.REAL_F64 if storage == _abi.STORE_F64 else _abi.REAL_F32
    h = O.HostLattice(X, Y, Z, collision=coll, storage=storage, pattern=_abi.PATTERN_AB, nu=NU, force=force,
                      macro_real=mreal, threads=2)
    rng = np.random.default_rng(seed)
    h.set_equilibrium_field(1 + 0.01 * rng.standard_normal((X, Y, Z)), *(0.03 * rng.standard_normal((3, X, Y, Z))))
    f = h.export_f64(0) * (1 + 0.01 * rng.standard_normal((27, X, Y, Z)))
    h.import_f64(f, which=0)
    return h, h.export_f64(0)


def pulled_momentum(f, x, y, z):
    p = np.array([f[q, (x - CX[q]) % X, (y - CY[q]) % Y, (z - CZ[q]) % Z] for q in range(27)])
    return (C * p[:, None]).sum(0)


@pytest.mark.parametrize("coll,factor", [(_abi.COLL_CUM, 1.0), (_abi.COLL_CUM_2017, 1.0), (_abi.COLL_MRT_LES, 0.5)])
def test_each_site_gains_its_own_force(coll, factor):
    """The momentum gained at each site is c F(site) with one constant c per
    operator: c = 1 for the cumulants (Geier's forcing); MRT_LES sees F only
    through u = (j + F/2)/rho in the rebuilt populations (col_mrt.h), so c =
    1/2.  A force plane read at the wrong site or component fails by O(|F|).
    (KBC is left out: its site-dependent stabiliser gamma also acts on a part
    that carries the F/2 momentum offset, so its gain is not a fixed multiple
    of F; the field reaches it through the same code, and the constant-field
    test below covers it bit for bit.)"""
    rng = np.random.default_rng(3)
    F = 1e-4 * rng.standard_normal((3, X, Y, Z))
    h, f = lattice(coll)
    h.set_force_field(*F)
    h.step()
    out = h.export_f64()
    gains = np.empty((3, X, Y, Z))
    for x in range(X):
        for y in range(Y):
            for z in range(Z):
                gains[:, x, y, z] = (C * out[:, x, y, z][:, None]).sum(0) - pulled_momentum(f, x, y, z)
                mass = sum(f[q, (x - CX[q]) % X, (y - CY[q]) % Y, (z - CZ[q]) % Z] for q in range(27))
                assert abs(out[:, x, y, z].sum() - mass) < 1e-15
    err = np.sqrt(((gains - factor * F) ** 2).sum(0)) / np.sqrt((F ** 2).sum(0))
    assert err.max() < 1e-11, (coll, err.max())


@pytest.mark.parametrize("storage", [_abi.STORE_F64, _abi.STORE_F32])
@pytest.mark.parametrize("coll", [_abi.COLL_CUM, _abi.COLL_KBC_N4])
def test_constant_field_equals_homogeneous_force(storage, coll):
    F0 = (2e-5, -3e-5, 1e-5)
    a, _ = lattice(coll, storage, force=F0, seed=5)
    b, _ = lattice(coll, storage, seed=5)
    b.set_force_field(*(np.full((X, Y, Z), c) for c in F0))
    for _ in range(3):
        a.step()
        b.step()
    assert np.array_equal(a.export_f64(), b.export_f64())


def test_field_adds_to_the_homogeneous_force():
    rng = np.random.default_rng(6)
    F = 1e-4 * rng.standard_normal((3, X, Y, Z))
    F0 = (1e-4, 2e-4, -1e-4)
    h, f = lattice(_abi.COLL_CUM, force=F0, seed=7)
    h.set_force_field(*F)
    h.step()
    out = h.export_f64()
    gain = (C * out[:, 2, 3, 1][:, None]).sum(0) - pulled_momentum(f, 2, 3, 1)
    assert np.abs(gain - (F[:, 2, 3, 1] + np.array(F0))).max() < 1e-16
