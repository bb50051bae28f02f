"""Cross-check of the reference's adjoint (sim_adjoint: the steady SRT adjoint
of d3q27/col_srt_adjoint.h, eq_adjoint.h, bc.h:295-398 and lbm_data.h:136-209)
at the level of its per-site operators, restated in oracle/sim_adjoint_ref.py:
each one against the exact transpose or derivative of the forward operator it
belongs to, by central differences of the restated forward (CPU, FP64).

- The adjoint collision is the exact transpose of the BGK collision's
  Jacobian at the primary run's rho, u (the same identity the product's own
  cumulant adjoint satisfies, tests/test_oracle_adjoint.py and the GPU tests).
- The data source added at the measured sites is the exact gradient of the
  misfit 1/2 (rho - rho_m)^2 + 1/2 |u - u_m|^2 with respect to the site's
  pre-collision populations.
- The inlet's gy is minus the exact u_y-derivative of the bounce-back inlet's
  replaced populations, contracted with their adjoint (a descent direction:
  the reference adds eps * gy to the profile, lbm_data.h:184).
- The inlet's "extra collision" is minus the adjoint of a density dependence
  of the replaced populations, taken with the adjoint before its collision;
  the forward inlet takes rho = 1 (the kernel struct's default), so the exact
  transpose of the forward has no such term: this is where the reference's
  adjoint departs from the discrete adjoint of its own forward step.
Each identity also fails for a planted mistake."""
import numpy as np
import pytest

import sim_adjoint_ref as R


def near_equilibrium(rng, amp=0.05):
    rho = 1.0 + 0.02 * rng.standard_normal()
    u = amp * rng.standard_normal(3)
    return R.feq(rho, u) * (1.0 + 0.03 * rng.standard_normal(27))


def central(fn, x, d, eps=1e-6):
    return (fn(x + eps * d) - fn(x - eps * d)) / (2 * eps)


@pytest.mark.parametrize("seed", range(6))
@pytest.mark.parametrize("nu", [0.0064, 0.02, 0.1])
def test_srt_adjoint_collision_is_the_transpose_of_bgk(seed, nu):
    rng = np.random.default_rng(seed)
    f = near_equilibrium(rng)
    rho, u = R.density_velocity(f)
    lam = rng.standard_normal(27)
    adj = R.srt_adjoint_collide(lam, rho, u, nu)
    for _ in range(4):
        d = rng.standard_normal(27)
        jd = central(lambda x: R.srt_collide(x, nu), f, d)
        assert abs(jd @ lam - d @ adj) <= 1e-8 * (np.abs(jd).sum() * np.abs(lam).max())


def test_a_wrong_adjoint_equilibrium_derivative_is_caught(monkeypatch):
    """Planted mistake: the velocity part of the adjoint equilibrium
    (feq_v, eq_adjoint.h:18-22) dropped."""
    def rho_only(k, q, u):
        ck = R.C[k]
        return 1.0 - 1.5 * (u @ u) + 3.0 * (ck @ u) + 4.5 * (ck @ u) ** 2
    monkeypatch.setattr(R, "_feq_adjoint", rho_only)
    rng = np.random.default_rng(3)
    f = near_equilibrium(rng)
    rho, u = R.density_velocity(f)
    lam = rng.standard_normal(27)
    adj = R.srt_adjoint_collide(lam, rho, u, 0.02)
    d = rng.standard_normal(27)
    jd = central(lambda x: R.srt_collide(x, 0.02), f, d)
    assert abs(jd @ lam - d @ adj) > 1e-4 * (np.abs(jd).sum() * np.abs(lam).max())


@pytest.mark.parametrize("seed", range(6))
def test_data_source_is_the_gradient_of_the_misfit(seed):
    rng = np.random.default_rng(10 + seed)
    f = near_equilibrium(rng)
    rho_m = 1.0 + 0.02 * rng.standard_normal()
    u_m = 0.05 * rng.standard_normal(3)

    def misfit(x):
        r, v = R.density_velocity(x)
        return 0.5 * (r - rho_m) ** 2 + 0.5 * ((v - u_m) @ (v - u_m))

    rho, u = R.density_velocity(f)
    src = R.data_source(rho, u, rho_m, u_m)
    for q in range(27):
        e = np.zeros(27)
        e[q] = 1.0
        g = central(misfit, f, e)
        assert abs(g - src[q]) <= 1e-9 * max(1.0, abs(src[q]))
    # planted mistake: the velocity misfit without the 1/rho of du/df
    wrong = np.array([(rho - rho_m) + (u - u_m) @ (R.C[q] - u) for q in range(27)])
    assert np.abs(wrong - src).max() > 1e-4 * np.abs(src).max()


@pytest.mark.parametrize("seed", range(6))
def test_inlet_gradient_is_minus_the_derivative_of_the_bounce_back_inlet(seed):
    rng = np.random.default_rng(20 + seed)
    own = near_equilibrium(rng)
    u_in = 0.04 * rng.standard_normal(3)
    lam = rng.standard_normal(27)  # the adjoint of the replaced +x populations (other entries unused)
    ey = np.array([0.0, 1.0, 0.0])
    du = central(lambda v: R.bb_inlet(own, v), u_in, ey)
    exact = du @ lam
    assert abs(R.inlet_gradient_y(lam, 1.0) + exact) <= 1e-9 * np.abs(lam).max()
    # planted mistake: the sign convention of the descent step flipped
    assert abs(-R.inlet_gradient_y(lam, 1.0) + exact) > 1e-3 * abs(exact)


@pytest.mark.parametrize("seed", range(6))
def test_inlet_extra_term_is_a_density_derivative_the_forward_does_not_have(seed):
    """With the adjoint loaded into the -x entries from the site's own +x
    slots (bc.h:331-339), inlet_extra = -sum_k lam_k d(bb_inlet_k)/d(rho)
    over the replaced directions k, evaluated at the adjoint before its
    collision.  The forward's inlet (rho = 1, independent of the
    populations) has d(bb_inlet)/d(f) = the own-slot mapping only, so the
    exact transpose subtracts nothing."""
    rng = np.random.default_rng(30 + seed)
    own = near_equilibrium(rng)
    u_in = 0.04 * rng.standard_normal(3)
    lam = rng.standard_normal(27)
    tmp = np.zeros(27)
    for k in R.PLUS_X:
        tmp[R.OPP[k]] = lam[k]  # KS.f[mzz] = slot pzz, ...
    drho = central(lambda r: R.bb_inlet(own, u_in, r), 1.0, 1.0)
    assert abs(R.inlet_extra(tmp, u_in) + drho @ lam) <= 1e-9 * np.abs(lam).max()
    # the forward inlet's derivative with respect to the populations: only the
    # own-slot mapping (unit entries), no density term
    for q in range(27):
        e = np.zeros(27)
        e[q] = 1.0
        dq = central(lambda x: R.bb_inlet(x, u_in), own, e)
        expect = np.zeros(27)
        for k in R.PLUS_X:
            if R.OPP[k] == q:
                expect[k] = 1.0
        assert np.abs(dq - expect).max() <= 1e-9
    assert abs(R.inlet_extra(tmp, u_in)) > 0  # the reference subtracts a non-zero term there
