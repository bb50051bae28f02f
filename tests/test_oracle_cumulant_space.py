"""Pins the oracle's three cumulant collisions in cumulant space.

The oracle computes a collision as the reference does (col_cum.h:13-511):
forward central-moment passes (Eqs 6-14), central moments -> cumulants (Eqs
51-54, 2017 variants), relaxation, cumulants -> central moments (Eqs 81-84),
backward passes (Eqs 88-96).  Most index and coefficient mistakes hide in
those transforms.  Here the cumulants of the populations before and after
one collision are computed by an independent route, the logarithm of the
discrete moment-generating function M(xi) = sum_q f_q exp(xi . (c_q - u)),
whose mixed derivatives of order <= 2 per axis are the D3Q27 cumulants
(Geier's C_abc = rho * kappa_abc).  That route is checked against mpmath's
numerical differentiation of log M and against the textbook fourth-order
joint-cumulant identity.  The post-collision cumulants must then obey the
relaxation rules the reference states in cumulant space:

  default (col_cum.h:247-259, 330-366): deviatoric second order * (1 - w1),
      trace -> rho (w2 = 1), every cumulant of order 3..6 -> 0;
  2017 (col_cum.h:193-297): the seven limited third-order combinations
      C* = (1 - w_n) C rho lambda / (rho lambda + |C|) with w3, w4, w5 the
      closed forms of Geier et al. 2017 (Part I) at w2 = 1, orders 4..6 -> 0;
  2017 + antialias (col_cum.h:234-245, 262-297): in addition the velocity
      derivatives D.. from the second-order cumulants enter the second-order
      Galilean correction, the fourth-order cumulants through A (Eqs 43-45,
      with the reference's Eq 44 D-term Dxu + Dyv - 2 Dzw as written) and the
      mixed ones through B (Eqs 46-48).

The closed forms are carried here expanded at w2 = 1 (integer coefficients
of polynomials in w1), not in the factored form the reference writes, so a
transcription mistake on either side shows.  Each rate is also read back from
the oracle's behaviour (the limiter's fixed ratio), so a rate on the wrong
channel fails too.  CPU only."""
import math

import mpmath
import numpy as np
import pytest

import oracle as O
from lbm_common import C

LAMBDA = 0.01  # col_cum.h:198-200


def poly(c, x):
    return sum(a * x ** (len(c) - 1 - i) for i, a in enumerate(c))


# Geier, Pasquali, Schoenherr 2017 (J. Comput. Phys. 348) closed forms at w2 = 1, expanded
def w3_of(w1):
    return poly([16, -24, -16], w1) / poly([7, -14, -8], w1)


def w4_of(w1):
    return poly([32, -120, 112], w1) / poly([9, -50, 56], w1)


def w5_of(w1):
    return poly([72, -312, 288, 96], w1) / poly([29, -130, 152, 48], w1)


def A_of(w1):
    return poly([-3, 2, 4], w1) / poly([5, -7, 2], w1)


def B_of(w1):
    return poly([-14, 28, 4], w1) / poly([15, -21, 6], w1)


# ---- cumulants from the log of the moment-generating function ---------------
def _mul(a, b):
    """Product of two series in (xi_x, xi_y, xi_z) truncated at degree 2 per axis."""
    out = np.zeros((3, 3, 3))
    for i in range(3):
        for j in range(3):
            for k in range(3):
                if a[i, j, k] != 0.0:
                    out[i:, j:, k:] += a[i, j, k] * b[:3 - i, :3 - j, :3 - k]
    return out


FACT = np.array([1.0, 1.0, 2.0])


def log_cumulants(f, u=None):
    """rho, u = j / rho, and K[a, b, c] = rho * (d/dxi)^(a,b,c) log M at xi = 0
    for a, b, c <= 2, with M(xi) = sum_q f_q exp(xi . (c_q - u)) / rho."""
    f = np.asarray(f, dtype=np.float64)
    rho = f.sum()
    if u is None:
        u = (C * f[:, None]).sum(axis=0) / rho
    d = C - np.asarray(u)[None, :]
    M = np.zeros((3, 3, 3))
    for a in range(3):
        for b in range(3):
            for c in range(3):
                M[a, b, c] = (f * d[:, 0] ** a * d[:, 1] ** b * d[:, 2] ** c).sum() / rho / (FACT[a] * FACT[b] * FACT[c])
    X = M.copy()
    X[0, 0, 0] = 0.0
    L = np.zeros((3, 3, 3))
    P = X.copy()
    for n in range(1, 7):  # X has no constant term: X^n starts at total degree n <= 6
        L += (-1) ** (n + 1) * P / n
        P = _mul(P, X)
    fa = FACT[:, None, None] * FACT[None, :, None] * FACT[None, None, :]
    return rho, np.asarray(u), rho * L * fa


def orders():
    return {(a, b, c): a + b + c for a in range(3) for b in range(3) for c in range(3)}


def test_log_cumulants_match_numerical_derivatives_of_log_mgf():
    """Every mixed derivative of log M, by mpmath at 30 digits."""
    rng = np.random.default_rng(11)
    f = O.equilibrium(1.02, 0.04, -0.03, 0.02, kind=0) * (1 + 0.05 * rng.standard_normal(27))
    rho, u, K = log_cumulants(f)
    mpmath.mp.dps = 30
    fm = [mpmath.mpf(float(x)) for x in f]
    dm = [[mpmath.mpf(int(C[q, a])) - mpmath.mpf(float(u[a])) for a in range(3)] for q in range(27)]

    def logM(x, y, z):
        return mpmath.log(sum(fm[q] * mpmath.exp(x * dm[q][0] + y * dm[q][1] + z * dm[q][2]) for q in range(27)))

    for (a, b, c) in [(1, 1, 0), (2, 0, 0), (1, 1, 1), (1, 2, 0), (2, 1, 1), (2, 2, 0), (1, 2, 2), (2, 2, 2)]:
        ref = float(mpmath.diff(logM, (0, 0, 0), (a, b, c))) * rho
        assert abs(K[a, b, c] - ref) < 1e-12 * max(1.0, abs(ref)), ((a, b, c), K[a, b, c], ref)


def test_log_cumulants_textbook_identities():
    """Zero-mean joint cumulants: k(X,X,Y,Y) = m_220 - m_200 m_020 - 2 m_110^2,
    k(X,X,Y,Z) = m_211 - m_200 m_011 - 2 m_110 m_101; a product-form
    distribution has no mixed cumulant; order <= 3 cumulants are central moments."""
    rng = np.random.default_rng(12)
    f = O.equilibrium(0.98, -0.05, 0.02, 0.06, kind=0) * (1 + 0.05 * rng.standard_normal(27))
    rho, u, K = log_cumulants(f)
    d = C - u[None, :]
    m = lambda a, b, c: (f * d[:, 0] ** a * d[:, 1] ** b * d[:, 2] ** c).sum() / rho  # noqa: E731
    assert abs(K[2, 2, 0] / rho - (m(2, 2, 0) - m(2, 0, 0) * m(0, 2, 0) - 2 * m(1, 1, 0) ** 2)) < 1e-15
    assert abs(K[2, 1, 1] / rho - (m(2, 1, 1) - m(2, 0, 0) * m(0, 1, 1) - 2 * m(1, 1, 0) * m(1, 0, 1))) < 1e-15
    for idx in [(1, 1, 0), (2, 0, 0), (1, 1, 1), (1, 2, 0), (0, 1, 2)]:
        assert abs(K[idx] / rho - m(*idx)) < 1e-15
    g = O.equilibrium(1.03, 0.05, -0.02, 0.07, kind=0)  # D3Q27_EQ_INV_CUM is a product of 1-D weights
    _, _, Kg = log_cumulants(g)
    for idx, n in orders().items():
        if n >= 2 and sum(1 for i in idx if i) >= 2:
            assert abs(Kg[idx]) < 1e-15, idx
    assert abs(Kg[2, 0, 0] - 1.03 / 3) < 1e-15


# ---- the reference's relaxation rules in cumulant space ----------------------
COMBOS = {  # limited third-order combinations: (+ index, - index or sign partner, sign, rate name)
    "120+102": ((1, 2, 0), (1, 0, 2), +1, "w3"), "210+012": ((2, 1, 0), (0, 1, 2), +1, "w3"),
    "201+021": ((2, 0, 1), (0, 2, 1), +1, "w3"), "120-102": ((1, 2, 0), (1, 0, 2), -1, "w4"),
    "210-012": ((2, 1, 0), (0, 1, 2), -1, "w4"), "201-021": ((2, 0, 1), (0, 2, 1), -1, "w4"),
}


def expected_post(K, rho, u, nu, coll):
    """Post-collision cumulants for the pre-collision cumulants K (F = 0)."""
    w1 = 1.0 / (3.0 * nu + 0.5)
    w2 = 1.0
    E = np.zeros((3, 3, 3))
    E[0, 0, 0] = K[0, 0, 0]
    ux2, uy2, uz2 = np.asarray(u) ** 2
    Dxu = Dyv = Dzw = DxvDyu = DxwDzu = DywDzv = 0.0
    if coll == 2:  # col_cum.h:237-245
        Dxu = (-w1 / (2 * rho) * (2 * K[2, 0, 0] - K[0, 2, 0] - K[0, 0, 2])
               - w2 / (2 * rho) * (K[2, 0, 0] + K[0, 2, 0] + K[0, 0, 2] - (rho - 1)))
        Dyv = Dxu + 1.5 * w1 / rho * (K[2, 0, 0] - K[0, 2, 0])
        Dzw = Dxu + 1.5 * w1 / rho * (K[2, 0, 0] - K[0, 0, 2])
        DxvDyu = -3 * w1 / rho * K[1, 1, 0]
        DxwDzu = -3 * w1 / rho * K[1, 0, 1]
        DywDzv = -3 * w1 / rho * K[0, 1, 1]
    # second order: Eqs 33-35
    e33 = (1 - w1) * (K[2, 0, 0] - K[0, 2, 0]) - 3 * rho * (1 - w1 / 2) * (ux2 * Dxu - uy2 * Dyv)
    e34 = (1 - w1) * (K[2, 0, 0] - K[0, 0, 2]) - 3 * rho * (1 - w1 / 2) * (ux2 * Dxu - uz2 * Dzw)
    e35 = (rho * w2 + (1 - w2) * (K[2, 0, 0] + K[0, 2, 0] + K[0, 0, 2])
           - 3 * rho * (1 - w2 / 2) * (ux2 * Dxu + uy2 * Dyv + uz2 * Dzw))
    E[2, 0, 0] = (e33 + e34 + e35) / 3
    E[0, 2, 0] = (-2 * e33 + e34 + e35) / 3
    E[0, 0, 2] = (e33 - 2 * e34 + e35) / 3
    for idx in [(1, 1, 0), (1, 0, 1), (0, 1, 1)]:
        E[idx] = (1 - w1) * K[idx]
    if coll == 0:
        return E
    rates = {"w3": w3_of(w1), "w4": w4_of(w1)}
    lim = {}
    for name, (p, m, s, r) in COMBOS.items():
        c = K[p] + s * K[m]
        lim[name] = (1 - rates[r]) * c * rho * LAMBDA / (rho * LAMBDA + abs(c))
    for a, b in [("120", "102"), ("210", "012"), ("201", "021")]:
        ia, ib = tuple(int(x) for x in a), tuple(int(x) for x in b)
        E[ia] = 0.5 * (lim[f"{a}+{b}"] + lim[f"{a}-{b}"])
        E[ib] = 0.5 * (lim[f"{a}+{b}"] - lim[f"{a}-{b}"])
    c111 = K[1, 1, 1]
    E[1, 1, 1] = (1 - w5_of(w1)) * c111 * rho * LAMBDA / (rho * LAMBDA + abs(c111))
    if coll == 2:
        A, B = A_of(w1), B_of(w1)
        v = 1 / w1 - 0.5
        e43 = 2 / 3 * v * A * rho * (Dxu - 2 * Dyv + Dzw)
        e44 = 2 / 3 * v * A * rho * (Dxu + Dyv - 2 * Dzw)  # as col_cum.h:282-283 writes it
        e45 = -4 / 3 * v * A * rho * (Dxu + Dyv + Dzw)
        E[2, 2, 0] = (e43 + e44 + e45) / 3
        E[2, 0, 2] = (-e43 + e45) / 3
        E[0, 2, 2] = (-e44 + e45) / 3
        E[2, 1, 1] = -v * B * rho * DywDzv / 3
        E[1, 2, 1] = -v * B * rho * DxwDzu / 3
        E[1, 1, 2] = -v * B * rho * DxvDyu / 3
    return E


def states(seed, n=12):
    rng = np.random.default_rng(seed)
    for _ in range(n):
        rho = 1 + 0.04 * rng.standard_normal()
        u = 0.06 * rng.standard_normal(3)
        f = O.equilibrium(rho, *u, kind=0) * (1 + 0.03 * rng.standard_normal(27))
        yield f


@pytest.mark.parametrize("coll", [0, 1, 2])
@pytest.mark.parametrize("nu", [0.002, 0.01, 0.03, 0.1])  # w1 = 1.99 ... 1.25, away from w4's pole at 14/9
def test_post_collision_cumulants_follow_the_relaxation_rules(coll, nu):
    for f in states(20 + coll):
        rho, u, K = log_cumulants(f)
        g, _ = O.collide(f, nu, coll)
        rho2, u2, K2 = log_cumulants(g, u)
        assert abs(rho2 - rho) < 1e-15 and np.abs((C * g[:, None]).sum(0) - rho * u).max() < 1e-15
        E = expected_post(K, rho, u, nu, coll)
        scale = max(1e-3, np.abs(K[1:, :, :]).max())
        for idx, n in orders().items():
            if n < 2:
                continue
            assert abs(K2[idx] - E[idx]) < 2e-13 * scale, (coll, nu, idx, K2[idx], E[idx], K[idx])


@pytest.mark.parametrize("coll", [1, 2])
def test_third_order_rates_are_the_closed_forms_per_channel(coll):
    """Read each channel's rate back from the oracle: with |C| << rho lambda the
    limiter is inactive to first order, so for a state at u = 0 carrying one
    combination C, C*/C -> 1 - w_n; the two amplitudes below extrapolate away
    the limiter's |C| / (rho lambda) term exactly (it is linear in |C|)."""
    rho = 1.05
    for nu in (0.003, 0.02, 0.1):
        w1 = 1 / (3 * nu + 0.5)
        expect = {"w3": w3_of(w1), "w4": w4_of(w1), "w5": w5_of(w1)}
        feq = O.equilibrium(rho, 0.0, 0.0, 0.0, kind=0)
        _, _, K0 = log_cumulants(feq)
        chans = dict(COMBOS)
        chans["111"] = ((1, 1, 1), None, 0, "w5")
        for name, (p, m, s, r) in chans.items():
            ratios = []
            for a in (1e-4, 2e-4):
                target = np.zeros((3, 3, 3))
                target[p] += a
                if m is not None:
                    target[m] += s * a
                h = moment_dual(target)
                g, _ = O.collide(feq + h, nu, coll)
                _, _, K2 = log_cumulants(g, (0.0, 0.0, 0.0))
                post = K2[p] + (s * K2[m] if m is not None else 0.0)
                pre = 2 * a if m is not None else a
                ratios.append(post / pre)
            # ratio(a) = (1 - w_n) rho lambda / (rho lambda + |C(a)|): invert, extrapolate |C| -> 0
            inv = [1 / x for x in ratios]
            inv0 = 2 * inv[0] - inv[1]
            w_n = 1 - 1 / inv0
            assert abs(w_n - expect[r]) < 1e-9 * max(1, abs(expect[r])), (coll, nu, name, w_n, expect[r])


def moment_dual(target):
    """Populations whose central moments at u = 0, sum_q h_q cx^a cy^b cz^c
    (a, b, c <= 2), equal target[a, b, c] (the D3Q27 moment matrix is invertible)."""
    E = np.array([[C[q, 0] ** a * C[q, 1] ** b * C[q, 2] ** c for q in range(27)]
                  for a in range(3) for b in range(3) for c in range(3)], dtype=float)
    return np.linalg.solve(E, target.reshape(27))


def test_closed_forms_match_the_reference_factored_forms_at_sample_points():
    """The expanded forms above against the factored forms evaluated with w2 = 1,
    and the published special values: every third-order rate vanishes at w1 = 2
    (the inviscid limit), and w4 has its pole at w1 = 14/9."""
    for w1 in (0.7, 1.0, 1.3, 1.7, 1.9, 1.99):
        w2 = 1.0
        w3f = 8 * (w1 - 2) * (w2 * (3 * w1 - 1) - 5 * w1) / (8 * (5 - 2 * w1) * w1 + w2 * (8 + w1 * (9 * w1 - 26)))
        w4f = 8 * (w1 - 2) * (w1 + w2 * (3 * w1 - 7)) / (w2 * (56 - 42 * w1 + 9 * w1 * w1) - 8 * w1)
        assert math.isclose(w3_of(w1), w3f, rel_tol=1e-13) and math.isclose(w4_of(w1), w4f, rel_tol=1e-13)
    for fn in (w3_of, w4_of, w5_of):
        assert abs(fn(2.0)) < 1e-15
    assert abs(poly([9, -50, 56], 14 / 9)) < 1e-12


# ---- linear dispersion of the shear wave -------------------------------------
def _shear_decay(J, kdir, udir, ks):
    """-log |lambda| of the shear mode of the linear step S(k) J (push
    streaming S(k) = diag exp(-i k . c_q)) for each wavenumber k."""
    kdir = np.asarray(kdir, float) / np.linalg.norm(kdir)
    udir = np.asarray(udir, float) / np.linalg.norm(udir)
    out = []
    for k in ks:
        lam, V = np.linalg.eig(np.diag(np.exp(-1j * k * (C @ kdir))) @ J)
        score = np.abs((udir @ C.T) @ V) / np.linalg.norm(V, axis=0) * (np.abs(lam) > 0.5)
        out.append(-np.log(np.abs(lam[np.argmax(score)])))
    return np.array(out)


def _k4_coefficient(J, kdir=(0, 1, 0), udir=(1, 0, 0)):
    """c4 of decay(k) = nu k^2 + c4 k^4 + c6 k^6 + ... (least squares over k <= 0.4)."""
    ks = np.linspace(0.05, 0.4, 12)
    s = _shear_decay(J, kdir, udir, ks)
    c = np.linalg.lstsq(np.vander(ks ** 2, 4, increasing=True), s / ks ** 2, rcond=None)[0]
    return c[0], c[1]


def _oracle_jacobian(coll, nu):
    """d collide / d f at the uniform rest state the operator maps to itself
    (for the antialias variant one collision away from the equilibrium),
    by central differences extrapolated to zero step: the limiter's |C| term
    makes the difference quotient a polynomial in the step."""
    f0, _ = O.collide(O.equilibrium(1.0, 0, 0, 0, kind=0), nu, coll)
    eps = np.array([1e-5, 2e-5, 4e-5, 8e-5])
    Js = []
    for e in eps:
        J = np.zeros((27, 27))
        for j in range(27):
            d = np.zeros(27)
            d[j] = e
            J[:, j] = (O.collide(f0 + d, nu, coll)[0] - O.collide(f0 - d, nu, coll)[0]) / (2 * e)
        Js.append(J.reshape(-1))
    return np.linalg.solve(np.vander(eps, 4, increasing=True), np.array(Js))[0].reshape(27, 27)


def test_shear_wave_dispersion_is_second_order_for_every_cumulant_variant():
    """Why a shear-wave convergence study shows no fourth-order rate in
    diffusive scaling: the decay per step of a linear shear wave of wavenumber
    k is nu k^2 + c4 k^4 + ..., and fourth-order accuracy needs c4 = 0.  For the
    reference's operators as written (w2 = 1, the closed-form rates, the Eq 44
    D-term as written) c4 is of order 1e-4..1e-3 for every variant.

    The analysis itself is checked on two-relaxation-time BGK, whose c4 for the
    same D3Q27 shear wave vanishes exactly on Lambda = 1/8 + (tau - 1/2)^2 / 2
    (tau the even rate, Lambda = (tau - 1/2)(tau_odd - 1/2)): the leading
    coefficient fits nu to 1e-9 and the zero to 1e-7."""
    from lbm_common import OPP, W
    Meq = W[:, None] * (1 + 3 * C @ C.T)
    I = np.eye(27)
    P = I[OPP]
    for tau in (0.53, 0.8):
        Lam = 0.125 + 0.5 * (tau - 0.5) ** 2
        tau_odd = 0.5 + Lam / (tau - 0.5)
        J = I - (0.5 / tau) * (I + P) @ (I - Meq) - (0.5 / tau_odd) * (I - P) @ (I - Meq)
        nu_fit, c4 = _k4_coefficient(J)
        assert abs(nu_fit - (tau - 0.5) / 3) < 1e-9 and abs(c4) < 1e-7, (tau, nu_fit, c4)
        J2 = I - (0.5 / tau) * (I + P) @ (I - Meq) - (0.5 / (tau_odd + 0.2)) * (I - P) @ (I - Meq)
        assert abs(_k4_coefficient(J2)[1]) > 1e-5
    nu = 0.01
    for coll in (0, 1, 2):
        nu_fit, c4 = _k4_coefficient(_oracle_jacobian(coll, nu))
        assert abs(nu_fit - nu) < 1e-7 * nu * 10, (coll, nu_fit)
        assert abs(c4) > 1e-4, (coll, c4)
