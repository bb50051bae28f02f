"""Pins the CPU oracle's single-site functions against what the mathematics
fixes: closed-form equilibrium moments, conservation, the equilibrium fixed
point, cube-symmetry equivariance, the deviation form's algebraic identity
and the 16-bit conversions.  The reference ships no golden vectors for this
path, so these checks are what make the oracle trustworthy."""
import numpy as np
import pytest
import torch

import oracle as O
from lbm_common import C, CX, CY, CZ, OPP, W, cube_symmetries, perm_of
from test_oracle_cumulant_space import w3_of, w4_of, w5_of

CUMS = [0, 1, 2]  # CUM, CUM_2017, CUM_2017_ANTIALIAS


def moments(f, u=(0.0, 0.0, 0.0)):
    c = C - np.asarray(u)[None, :]
    return c, f


@pytest.mark.parametrize("kind", [0, 1])
@pytest.mark.parametrize("rho,u", [(1.0, (0.0, 0.0, 0.0)), (1.03, (0.05, -0.02, 0.07)), (0.97, (-0.11, 0.0, 0.03))])
def test_equilibrium_moments(kind, rho, u):
    f = O.equilibrium(rho, *u, kind=kind)
    assert abs(f.sum() - rho) < 1e-15
    for a in range(3):
        assert abs((C[:, a] * f).sum() - rho * u[a]) < 1e-15
        for b in range(3):
            expect = rho * (u[a] * u[b] + (1.0 / 3.0 if a == b else 0.0))
            assert abs((C[:, a] * C[:, b] * f).sum() - expect) < 1e-15


@pytest.mark.parametrize("rho,u", [(1.0, (0.0, 0.0, 0.0)), (1.03, (0.05, -0.02, 0.07))])
def test_product_equilibrium_factorises(rho, u):
    """D3Q27_EQ_INV_CUM is rho * prod_d W(c_d; u_d): the weights factorise, so
    every central moment is the product of 1-D central moments (kappa_2 = 1/3)."""
    f = O.equilibrium(rho, *u, kind=0)
    d = C - np.asarray(u)[None, :]
    for (i, j, k) in [(2, 2, 0), (2, 0, 2), (0, 2, 2), (2, 2, 2), (1, 1, 0), (2, 1, 0), (1, 2, 1)]:
        m = (d[:, 0] ** i * d[:, 1] ** j * d[:, 2] ** k * f).sum()
        k1 = [1.0, 0.0, 1.0 / 3.0]
        assert abs(m - rho * k1[i] * k1[j] * k1[k]) < 1e-15


def near_eq(rng, rho=1.0, u=(0.03, -0.02, 0.01), eps=0.01):
    f = O.equilibrium(rho, *u, kind=0)
    return f * (1 + eps * rng.standard_normal(27))


@pytest.mark.parametrize("coll", CUMS)
def test_collision_conserves_mass_and_momentum(coll):
    rng = np.random.default_rng(1)
    for _ in range(20):
        f = near_eq(rng)
        g, m = O.collide(f, 0.02 / 3, coll)
        assert abs(g.sum() - f.sum()) < 1e-15
        for a in range(3):
            assert abs((C[:, a] * g).sum() - (C[:, a] * f).sum()) < 1e-15


@pytest.mark.parametrize("coll", CUMS)
def test_forcing_adds_F_to_momentum(coll):
    """Geier 2017 forcing (col_cum.h:371-373): the collision adds F to sum c f."""
    rng = np.random.default_rng(2)
    F = (1e-4, -2e-4, 3e-5)
    f = near_eq(rng)
    g, _ = O.collide(f, 0.01, coll, F=F)
    for a in range(3):
        assert abs((C[:, a] * g).sum() - (C[:, a] * f).sum() - F[a]) < 1e-15


@pytest.mark.parametrize("coll", [0, 1])
def test_equilibrium_is_fixed_point(coll):
    for rho, u in [(1.0, (0, 0, 0)), (1.02, (0.03, -0.02, 0.01)), (0.95, (0.1, 0.08, -0.05))]:
        f = O.equilibrium(rho, *u, kind=0)
        g, _ = O.collide(f, 0.02 / 3, coll)
        assert np.abs(g - f).max() < 5e-16 * 8


def test_antialias_moves_rest_state_only_in_fourth_order():
    """With USE_GEIER_CUM_ANTIALIAS the trace term subtracts rho-1 (col_cum.h:239),
    so even the rest equilibrium is relaxed: at u = 0 only the 4th-order
    cumulants C_220 = C_202 = C_022 change, isotropically."""
    f = O.equilibrium(1.0, 0, 0, 0, kind=0)
    g, _ = O.collide(f, 0.02 / 3, 2)
    d = g - f
    assert abs(d.sum()) < 1e-15
    for a in range(3):
        assert abs((C[:, a] * d).sum()) < 1e-15
        for b in range(3):
            assert abs((C[:, a] * C[:, b] * d).sum()) < 1e-15  # second moments unchanged
    m220 = (CX**2 * CY**2 * d).sum()
    m202 = (CX**2 * CZ**2 * d).sum()
    m022 = (CY**2 * CZ**2 * d).sum()
    assert abs(m220) > 1e-4 and abs(m220) < 1e-1
    assert abs(m220 - m202) < 1e-15 and abs(m220 - m022) < 1e-15
    for M in cube_symmetries():
        P = perm_of(M)
        assert np.abs(g[P] - g).max() < 1e-15


@pytest.mark.parametrize("coll", CUMS)
def test_cube_symmetry_equivariance(coll):
    """collide(R f) = R collide(f) for all 48 cube symmetries: catches index typos.

    The antialiased 2017 variant is only reflection-equivariant: its Eq 43-45
    right-hand sides (col_cum.h:280-284) give C*_220 the term -(2/3) visc A rho
    (Dyv + Dzw) and C*_022 the term (Dxu + Dyv), i.e. the velocity-gradient
    terms of the xy and yz cumulants are exchanged, which breaks the axis
    permutations.  The oracle reproduces the reference as written."""
    rng = np.random.default_rng(3)
    f = near_eq(rng, rho=1.01, u=(0.04, -0.03, 0.02), eps=0.02)
    g, _ = O.collide(f, 0.01, coll)
    group = cube_symmetries()
    if coll == 2:
        group = [M for M in group if (np.abs(M) == np.eye(3, dtype=int)).all()]
        assert len(group) == 8
        # and the permutations really are broken (the quirk is present)
        M = np.array([[0, 1, 0], [1, 0, 0], [0, 0, 1]])
        P = perm_of(M)
        fr = np.empty(27)
        fr[P] = f
        gr, _ = O.collide(fr, 0.01, coll)
        expect = np.empty(27)
        expect[P] = g
        assert np.abs(gr - expect).max() > 1e-8
    for M in group:
        P = perm_of(M)  # direction q maps to P[q]
        fr = np.empty(27)
        fr[P] = f
        gr, _ = O.collide(fr, 0.01, coll)
        expect = np.empty(27)
        expect[P] = g
        assert np.abs(gr - expect).max() < 2e-15


def test_deviation_form_equals_absolute_in_double():
    """The deviation (well-conditioned) algebra is an identity: evaluated in
    double it reproduces the absolute cumulant to rounding."""
    rng = np.random.default_rng(4)
    for F in [(0, 0, 0), (1e-4, 2e-5, -3e-5)]:
        for _ in range(10):
            f = near_eq(rng, rho=1 + 0.02 * rng.standard_normal(), u=0.05 * rng.standard_normal(3), eps=0.02)
            g_abs, _ = O.collide(f, 0.003, 0, F=F)
            g_dev = O.collide_dev(f - W, 0.003, F=F, single=False)
            assert np.abs(g_dev + W - g_abs).max() < 1e-15


def test_deviation_form_float_accuracy():
    rng = np.random.default_rng(5)
    f = near_eq(rng, eps=0.02)
    g_abs, _ = O.collide(f, 0.003, 0)
    g32 = O.collide_dev(f - W, 0.003, single=True)
    assert np.abs((g32 + W) - g_abs).max() < 1e-7


def test_f16_conversion_matches_ieee():
    rng = np.random.default_rng(6)
    vals = np.concatenate([rng.standard_normal(20000) * 10.0 ** rng.integers(-9, 5, 20000),
                           np.array([0.0, -0.0, 65504.0, 65520.0, 1e-8, 2.98e-8, 6.1e-5, 6.0e-5])]).astype(np.float32)
    L = O.lib()
    for v in vals:
        h = L.oracle_f32_to_f16(float(v))
        ref = np.array([v], dtype=np.float32).astype(np.float16).view(np.uint16)[0]
        assert h == ref, (v, h, ref)
        back = L.oracle_f16_to_f32(h)
        assert np.float32(back) == np.array([ref], dtype=np.uint16).view(np.float16).astype(np.float32)[0] or np.isnan(back)


def test_bf16_conversion_matches_torch():
    rng = np.random.default_rng(7)
    vals = (rng.standard_normal(20000) * 10.0 ** rng.integers(-30, 30, 20000)).astype(np.float32)
    ref = torch.from_numpy(vals).to(torch.bfloat16).view(torch.int16).numpy().view(np.uint16)
    L = O.lib()
    for v, r in zip(vals, ref):
        assert L.oracle_f32_to_bf16(float(v)) == r
    for r in ref[:200]:
        assert np.float32(L.oracle_bf16_to_f32(int(r))) == np.array([int(r) << 16], dtype=np.uint32).view(np.float32)[0]


def test_snorm16_conversion_is_round_to_nearest_even():
    """The fixed-point deviation form stores round(32767 * clamp(x, -1, 1))."""
    import ctypes as C
    L = O.lib()
    rng = np.random.default_rng(8)
    x = np.concatenate([rng.uniform(-1.2, 1.2, 5000), (np.arange(-20, 20) + 0.5) / 32767.0]).astype(np.float32)
    lat = O.HostLattice(1, 1, 1, storage=5, store_range=1.0)
    # encode through an import of absolute values f = w (1 + x) for the rest site
    from lbm_common import W
    for chunk in np.array_split(x, 200):
        for v in chunk[:3]:
            f = (W * (1.0 + np.float64(v))).reshape(27, 1, 1, 1)
            lat.import_f64(f)
            got = lat.df[0].reshape(27)[0]
            want = np.int16(np.rint(np.float32(np.clip(np.float32(np.float32(f[0, 0, 0, 0] - W[0]) * np.float32((np.float32(1.0 / W[0])) * np.float32(1.0))), -1, 1)) * np.float32(32767.0)))
            assert got == want


def test_i16_class_ranges_round_trip():
    """With one bound R_n per direction class, an import/export round trip of
    the fixed-point form is exact to half a quantum w_q R_n / 32767 of its
    class, and a deviation beyond the class bound saturates at it."""
    from lbm_common import CX, CY, CZ, W
    Rc = (0.01, 0.2, 0.3, 0.4)
    cls = np.array([int(c[0] != 0) + int(c[1] != 0) + int(c[2] != 0) for c in zip(CX, CY, CZ)])
    R = np.asarray(Rc)[cls]
    lat = O.HostLattice(4, 3, 2, storage=5, store_range=1.0, store_range_class=Rc)
    rng = np.random.default_rng(2)
    x = rng.uniform(-1, 1, (27, 4, 3, 2)) * R[:, None, None, None] * 0.999
    f = W[:, None, None, None] * (1 + x)
    lat.import_f64(f)
    back = lat.export_f64()
    q = (W * R / 32767)[:, None, None, None]
    assert (np.abs(back - f) <= 0.5 * q * 1.0001 + 1e-15).all()
    f2 = W[:, None, None, None] * (1 + 2 * R[:, None, None, None] * np.sign(x))
    lat.import_f64(f2)
    sat = lat.export_f64() - W[:, None, None, None]
    assert np.allclose(np.abs(sat), (W * R)[:, None, None, None], rtol=1e-6)


def moment_dual(i, j, k):
    """Populations whose raw moments sum_q h_q cx^a cy^b cz^c (a, b, c <= 2)
    are 1 for (i, j, k) and 0 for the other 26 (the D3Q27 moment matrix is
    invertible)."""
    E = np.array([[CX[q] ** a * CY[q] ** b * CZ[q] ** c for q in range(27)]
                  for a in range(3) for b in range(3) for c in range(3)], dtype=float)
    e = np.zeros(27)
    e[9 * i + 3 * j + k] = 1.0
    return np.linalg.solve(E, e)


# third-order combinations with their own limiter (col_cum.h:198-227) and the
# closed-form rate of each channel (Geier et al. 2017 at w2 = 1)
LIMITED = {"120+102": (((1, 2, 0), 0.5), ((1, 0, 2), 0.5)), "210-012": (((2, 1, 0), 0.5), ((0, 1, 2), -0.5)),
           "111": (((1, 1, 1), 1.0),)}
CHANNEL_RATE = {"120+102": w3_of, "210-012": w4_of, "111": w5_of}


@pytest.mark.parametrize("coll", [1, 2])
@pytest.mark.parametrize("combo", list(LIMITED))
def test_2017_limiter_form_and_lambda(coll, combo):
    """The 2017 variant relaxes each limited third-order cumulant combination
    C with omega = omega_n + (1 - omega_n) |C| / (rho lambda + |C|) (Geier et
    al. 2017), so C* / C = (1 - omega_n) rho lambda / (rho lambda + |C|): its
    inverse is linear in |C| with intercept / (slope rho) = lambda = 0.01.
    At u = 0 a perturbation carrying only this combination leaves the other
    cumulants as they were, and order-3 cumulants are the raw moments.  rho
    differs from 1 so that rho lambda and lambda are told apart.  The intercept
    gives the channel's rate w_n, which must be the closed form for that
    channel (w3 for C_120 + C_102, w4 for C_210 - C_012, w5 for C_111)."""
    rho, nu = 1.05, 0.01
    feq = O.equilibrium(rho, 0.0, 0.0, 0.0, kind=0)
    h = sum(wt * moment_dual(*m) for m, wt in LIMITED[combo])
    sign = {m: (1.0 if wt > 0 else -1.0) for m, wt in LIMITED[combo]}

    def combo_of(f):
        return sum(sign[m] * (CX ** m[0] * CY ** m[1] * CZ ** m[2] * f).sum() for m in sign)

    amps = np.array([2e-4, 2e-3, 1e-2, 4e-2])
    inv = []
    for a in amps:
        f = feq + a * h
        assert abs(combo_of(f) - a) < 1e-15
        g, _ = O.collide(f, nu, coll)
        inv.append(a / combo_of(g))
    inv = np.array(inv)
    S, I = np.polyfit(amps, inv, 1)
    assert np.abs(I + S * amps - inv).max() < 1e-9 * np.abs(inv).max()
    assert abs(I / (S * rho) / 0.01 - 1) < 1e-8
    omega_n = 1 - 1 / I
    expect = CHANNEL_RATE[combo](1 / (3 * nu + 0.5))
    assert abs(omega_n - expect) < 1e-8, (omega_n, expect)
