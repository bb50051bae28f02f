"""Pins of the oracle's alternative D3Q27 operators: MRT_LES (col_mrt.h),
KBC N1-N4 / C1-C4 (col_kbc_n.h, col_kbc_c.h) and the entropic equilibrium
(eq_entropic.h) -- against closed forms and invariants, not against a
retyped formula:

- the entropic equilibrium carries exactly rho and rho u;
- every operator conserves mass and momentum (F = 0);
- MRT_LES keeps the polynomial equilibrium fixed, relaxes all six
  non-equilibrium components of Pi by one common factor (1 - omega), and that
  factor is the Smagorinsky rate of Hou et al. (1996),
  tau_eff = (tau + sqrt(tau^2 + 18 C sqrt(Q) / rho)) / 2, tending to 1/tau;
- every operator is equivariant under the 48 cube symmetries (index and
  sign slips in the 27 per-direction expressions break this);
- a small shear wave decays as exp(-nu k^2 t) (KBC: nu = cs^2/2 (1/beta - 1));
- the KBC stabiliser gamma is the stationary point of the quadratic entropy
  of the post-collision state (Karlin, Boesch, Chikatamarla 2014).
"""
import numpy as np
import pytest

import oracle as O
from lbm_common import C as CV, cube_symmetries, perm_of
from pytnl_lbm import _abi

MRT = _abi.COLL_MRT_LES
KBC = {"N1": _abi.COLL_KBC_N1, "N2": _abi.COLL_KBC_N2, "N3": _abi.COLL_KBC_N3, "N4": _abi.COLL_KBC_N4,
       "C1": _abi.COLL_KBC_C1, "C2": _abi.COLL_KBC_C2, "C3": _abi.COLL_KBC_C3, "C4": _abi.COLL_KBC_C4}
ALT = {"MRT": MRT, **KBC}


def near_eq(rng, rho=1.0, u=(0.0, 0.0, 0.0), eps=0.01):
    f = O.equilibrium(rho, *u, kind=1)
    return f * (1 + eps * rng.standard_normal(27))


@pytest.mark.parametrize("u", [(0.0, 0.0, 0.0), (0.05, -0.02, 0.03), (0.1, 0.08, -0.12)])
def test_entropic_equilibrium_moments(u):
    rho = 1.07
    f = O.equilibrium(rho, *u, kind=2)
    assert abs(f.sum() / rho - 1) < 1e-14
    np.testing.assert_allclose(CV.T @ f, rho * np.array(u), atol=1e-15)
    if u == (0.0, 0.0, 0.0):
        np.testing.assert_allclose(f, O.equilibrium(rho, 0, 0, 0, kind=1), rtol=1e-14)


@pytest.mark.parametrize("name", list(ALT))
def test_conserves_mass_and_momentum(name):
    rng = np.random.default_rng(11)
    f = near_eq(rng, 1.02, (0.03, -0.02, 0.04), 0.02)
    g, _ = O.collide(f, 0.01, ALT[name])
    assert abs(g.sum() / f.sum() - 1) < 1e-14
    np.testing.assert_allclose(CV.T @ g, CV.T @ f, atol=2e-16)


def second_moments(f):
    return np.array([(CV[:, a] * CV[:, b] * f).sum() for a, b in ((0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2))])


def test_mrt_keeps_its_equilibrium():
    f = O.equilibrium(1.03, 0.04, -0.03, 0.02, kind=1)
    g, _ = O.collide(f, 0.02, MRT)
    np.testing.assert_allclose(g, f, rtol=2e-15, atol=1e-17)


@pytest.mark.parametrize("eps", [1e-7, 1e-3, 3e-2])
def test_mrt_relaxes_pi_with_the_smagorinsky_rate(eps):
    rng = np.random.default_rng(2)
    nu, rho, u = 0.01, 1.02, np.array([0.03, 0.01, -0.02])
    f = O.equilibrium(rho, *u, kind=1) * (1 + eps * rng.standard_normal(27))
    g, mac = O.collide(f, nu, MRT)
    rho, u = mac[0], mac[1:]
    peq = rho * np.array([u[0] ** 2 + 1 / 3, u[1] ** 2 + 1 / 3, u[2] ** 2 + 1 / 3, u[0] * u[1], u[0] * u[2], u[1] * u[2]])
    pin, pout = second_moments(f) - peq, second_moments(g) - peq
    ratio = pout / pin
    assert np.ptp(ratio) < 1e-9 * max(1.0, 1 / eps)
    Q = 2 * (pin[:3] @ pin[:3] + 2 * pin[3:] @ pin[3:])
    tau = 3 * nu + 0.5
    tau_eff = 0.5 * (tau + np.sqrt(tau * tau + 18 * 0.0342 * np.sqrt(Q) / rho))
    assert abs(ratio.mean() - (1 - 1 / tau_eff)) < 1e-9 * max(1.0, 1 / eps)
    if eps < 1e-6:
        assert abs(ratio.mean() - (1 - 1 / tau)) < 1e-5


@pytest.mark.parametrize("name", list(ALT))
def test_cube_symmetry_equivariance(name):
    rng = np.random.default_rng(3)
    f = near_eq(rng, 1.01, (0.04, -0.03, 0.02), 0.02)
    g, _ = O.collide(f, 0.01, ALT[name])
    for M in cube_symmetries():
        P = perm_of(M)
        fr = np.empty(27)
        fr[P] = f
        gr, _ = O.collide(fr, 0.01, ALT[name])
        expect = np.empty(27)
        expect[P] = g
        assert np.abs(gr - expect).max() < 1e-15, M


@pytest.mark.parametrize("name", ["MRT", "N1", "N4", "C1", "C4"])
def test_shear_wave_decay_rate(name):
    X, Y, Z = 2, 64, 2
    nu = 0.05
    U = 1e-4
    y = np.arange(Y)
    k = 2 * np.pi / Y
    ux = np.broadcast_to(U * np.sin(k * y)[None, :, None], (X, Y, Z)).copy()
    lat = O.HostLattice(X, Y, Z, collision=ALT[name], nu=nu, threads=4)
    lat.set_equilibrium_field(np.ones((X, Y, Z)), ux, 0 * ux, 0 * ux)
    amps = []
    steps = 400
    for _ in range(steps):
        lat.step(output_macro=True)
        vx = lat.macros()[1][0, :, 0]
        amps.append(2 / Y * (vx * np.sin(k * y)).sum())
    t = np.arange(steps)
    rate = -np.polyfit(t[50:], np.log(np.array(amps)[50:]), 1)[0]
    assert abs(rate / (nu * k * k) - 1) < 5e-3, rate / (nu * k * k)


@pytest.mark.parametrize("name,kind", [("MRT", 1), ("N1", 1), ("C1", 1), ("N2", 2), ("N4", 2), ("C3", 2)])
def test_lattice_initialises_with_the_operators_equilibrium(name, kind):
    lat = O.HostLattice(3, 4, 5, collision=ALT[name], nu=0.01)
    lat.set_equilibrium(1.02, 0.05, -0.04, 0.03)
    np.testing.assert_allclose(lat.export_f64()[:, 1, 2, 3], O.equilibrium(1.02, 0.05, -0.04, 0.03, kind=kind),
                               rtol=1e-15)


def product_equilibrium(rho, u):
    """Tensor product of the D1Q3 equilibria (moments 1, u, 1/3 + u^2 per
    axis), the equilibrium of the KBC operators (col_kbc_n.h:289-330)."""
    def phi(c, v):
        return np.where(c == 0, 2.0 / 3.0 - v * v, (1.0 / 3.0 + v * v + c * v) / 2.0)
    return rho * phi(CV[:, 0], u[0]) * phi(CV[:, 1], u[1]) * phi(CV[:, 2], u[2])


def test_product_equilibrium_moments():
    rho, u = 1.03, np.array([0.07, -0.04, 0.05])
    feq = product_equilibrium(rho, u)
    assert abs(feq.sum() - rho) < 1e-15
    np.testing.assert_allclose(CV.T @ feq, rho * u, atol=1e-16)
    np.testing.assert_allclose([(CV[:, d] ** 2 * feq).sum() for d in range(3)], rho * (1 / 3 + u * u), rtol=1e-14)


@pytest.mark.parametrize("seed", [1, 2, 3])
@pytest.mark.parametrize("name", list(KBC))
def test_kbc_stabiliser_is_the_entropic_stationary_point(name, seed):
    """KBC: f* = f - beta (2 ds + gamma dh) with f - feq = ds + dh, and gamma
    chosen where the quadratic entropy sum (f* - feq)^2 / feq is stationary in
    gamma, i.e. <dh | f* - feq> = 0 (weights 1/feq).  With a = f - feq and
    b = f - f* = beta (2 a + (gamma - 2) dh), dh is parallel to
    r = b / beta - 2 a for any split of a into ds and dh, so the output must
    satisfy <r | f* - feq> = 0 with beta = 1 / (2 nu / cs^2 + 1).  This pins the
    stabiliser and beta; r must not vanish (gamma = 2 is plain BGK, for which
    the identity holds trivially)."""
    rng = np.random.default_rng(seed)
    nu = 0.01
    f = near_eq(rng, 1.02, (0.05, -0.03, 0.04), 0.03)
    rho = f.sum()
    u = CV.T @ f / rho
    g, _ = O.collide(f, nu, KBC[name])
    feq = product_equilibrium(rho, u)
    beta = 1.0 / (2.0 * nu * 3.0 + 1.0)
    a, b = f - feq, f - g
    r = b / beta - 2.0 * a

    def ip(x, y):
        return float((x * y / feq).sum())

    assert np.sqrt(ip(r, r)) > 1e-2 * np.sqrt(ip(a, a))
    assert abs(ip(r, g - feq)) <= 1e-10 * np.sqrt(ip(r, r) * ip(g - feq, g - feq))
