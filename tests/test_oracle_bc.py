"""Pins of the oracle's boundary codes that the parity tests leave open,
each by the property its reference lines define (d3q27/bc.h, d3q7/bc.h),
checked on one A-B step from a seeded random state: the step pulls
f_q(x - c_q), applies the code, collides where doCollision says (bc.h:406-412)
and stores the result at x, so the stored populations of a boundary site
are a closed-form function of the state the step read.

- SYM_* (bc.h:222-293): not collided; the stored populations are the pulled
  ones with the incoming set replaced by their mirror images.
- INFLOW_LEFT (bc.h:91-142, the moment condition): the reconstructed
  populations carry rho = (f_0-plane + 2 f_-x-plane) / (1 - u_x), j = rho u_in
  and Pi_yz = rho u_y u_z; the default cumulant conserves rho and j and
  relaxes the central moment k_011 (zero here) to zero, so all three survive
  the collision.
- INFLOW_BB_LEFT (bc.h:143-165): the nine +x populations are the site's own
  unstreamed opposite slots minus 6 w_q rho c_q.u (rho = 1); the collision
  conserves the resulting rho and j.
- OUTFLOW_RIGHT_INTERP (bc.h:195-199, streaming_AB.h:210-244): the -x
  populations interpolate cs f(x-1) + (1-cs) f(x), then f += feq(1,u) -
  feq(rho,u); after the collision j = j'/rho' of the interpolated set.
- D3Q7 INFLOW and SYM_* (d3q7/bc.h): not collided; feq(phi_in, u) of the
  NSE velocity, and mirrored pulls.
- D3Q7 OUTFLOW_RIGHT (d3q7/bc.h:73-92, 202): every population is pulled from
  the column x - 1 and the site then collides like a fluid site.
"""
import numpy as np
import pytest

import oracle as O
from lbm_common import CX, CY, CZ, C, W
from pytnl_lbm import _abi
from pytnl_lbm._abi import ADE_GEO, GEO

X, Y, Z = 8, 7, 9
NU = 0.02


def qidx(c):
    return int(np.where((C == np.asarray(c)).all(axis=1))[0][0])


def lattice(geo, seed=1, u_scale=0.03):
    h = O.HostLattice(X, Y, Z, storage=_abi.STORE_F64, pattern=_abi.PATTERN_AB, nu=NU, geo=geo, threads=4)
    h.set_inflow(0.04, 0.01, -0.015)
    rng = np.random.default_rng(seed)
    h.set_equilibrium_field(1 + 0.01 * rng.standard_normal((X, Y, Z)), *(u_scale * rng.standard_normal((3, X, Y, Z))))
    f = h.export_f64(0) * (1 + 0.02 * rng.standard_normal((27, X, Y, Z)))
    h.import_f64(f, which=0)
    return h, h.export_f64(0)


def pulled(f, x, y, z):
    return np.array([f[q, x - CX[q], y - CY[q], z - CZ[q]] for q in range(27)])


def moments(g):
    return g.sum(), np.array([(CX * g).sum(), (CY * g).sum(), (CZ * g).sum()]), (CY * CZ * g).sum()


def interior_plane(code, axis, pos):
    geo = np.full((X, Y, Z), GEO.PERIODIC, dtype=np.int16)
    sl = [slice(None)] * 3
    sl[axis] = pos
    geo[tuple(sl)] = code
    return geo


SYM = {"SYM_TOP": (2, -1), "SYM_BOTTOM": (2, 1), "SYM_LEFT": (0, 1), "SYM_RIGHT": (0, -1),
       "SYM_BACK": (1, 1), "SYM_FRONT": (1, -1)}


@pytest.mark.parametrize("name", list(SYM))
def test_symmetry_planes_mirror_the_incoming_populations(name):
    """SYM_TOP sets the populations with c_z = -1 to their z mirrors
    (bc.h:222-232), SYM_BOTTOM those with c_z = +1, and so on."""
    axis, sgn = SYM[name]
    pos = 4
    h, f = lattice(interior_plane(getattr(GEO, name), axis, pos), seed=2)
    h.step()
    out = h.export_f64()
    for a in range(1, X - 1) if axis != 0 else [pos]:
        for b in range(1, Y - 1) if axis != 1 else [pos]:
            for c in range(1, Z - 1) if axis != 2 else [pos]:
                p = pulled(f, a, b, c)
                want = p.copy()
                for q in range(27):
                    cq = C[q].copy()
                    if cq[axis] == sgn:
                        cq[axis] = -sgn
                        want[q] = p[qidx(cq)]
                assert np.array_equal(out[:, a, b, c], want), (name, a, b, c)


def inflow_left_geo():
    geo = np.full((X, Y, Z), GEO.PERIODIC, dtype=np.int16)
    geo[3] = GEO.INFLOW_LEFT
    return geo


def test_moment_inflow_prescribes_rho_j_and_pi_yz():
    h, f = lattice(inflow_left_geo(), seed=3)
    u = np.array([0.04, 0.01, -0.015])
    h.step()
    out = h.export_f64()
    for y in range(1, Y - 1):
        for z in range(1, Z - 1):
            p = pulled(f, 3, y, z)
            zero = [q for q in range(27) if CX[q] == 0]
            minus = [q for q in range(27) if CX[q] == -1]
            rho = (p[zero].sum() + 2 * p[minus].sum()) / (1 - u[0])
            r, j, pyz = moments(out[:, 3, y, z])
            assert abs(r - rho) < 1e-14
            assert np.abs(j - rho * u).max() < 1e-14
            assert abs(pyz - rho * u[1] * u[2]) < 1e-14


def test_bounce_back_inflow_replaces_the_plus_x_populations():
    geo = np.full((X, Y, Z), GEO.PERIODIC, dtype=np.int16)
    geo[3] = GEO.INFLOW_BB_LEFT
    h, f = lattice(geo, seed=4)
    u = np.array([0.04, 0.01, -0.015])
    h.step()
    out = h.export_f64()
    for y in range(1, Y - 1):
        for z in range(1, Z - 1):
            g = pulled(f, 3, y, z)
            for q in range(27):
                if CX[q] == 1:
                    opp = qidx(-C[q])
                    g[q] = f[opp, 3, y, z] - 2 * W[q] * 1.0 * 3 * (-(C[q] @ u))
            r, j, _ = moments(g)
            r2, j2, _ = moments(out[:, 3, y, z])
            assert abs(r2 - r) < 1e-14
            assert np.abs(j2 - j).max() < 1e-14


def test_interpolating_outflow_carries_the_interpolated_velocity():
    geo = np.full((X, Y, Z), GEO.PERIODIC, dtype=np.int16)
    geo[5] = GEO.OUTFLOW_RIGHT_INTERP
    h, f = lattice(geo, seed=5)
    cs = 0.5773502691896257
    h.step()
    out = h.export_f64()
    for y in range(1, Y - 1):
        for z in range(1, Z - 1):
            g = np.empty(27)
            for q in range(27):
                src = lambda xx: f[q, xx, y - CY[q], z - CZ[q]]  # noqa: E731
                if CX[q] < 0:
                    g[q] = cs * src(4) + (1 - cs) * src(5)
                else:
                    g[q] = src(4 if CX[q] > 0 else 5)
            r, j, _ = moments(g)
            r2, j2, _ = moments(out[:, 5, y, z])
            assert abs(r2 - 1.0) < 1e-14
            assert np.abs(j2 - j / r).max() < 1e-14


def test_d3q7_inflow_and_symmetry():
    """D3Q7 INFLOW writes feq(phi_in, u) with the NSE site's velocity, SYM_TOP
    mirrors the z populations; neither collides (d3q7/bc.h)."""
    geo = np.full((X, Y, Z), GEO.PERIODIC, dtype=np.int16)
    ageo = np.full((X, Y, Z), ADE_GEO.PERIODIC, dtype=np.int16)
    ageo[3] = ADE_GEO.INFLOW
    ageo[:, :, 6] = ADE_GEO.SYM_TOP
    h, _ = lattice(geo, seed=6)
    a = O.HostAde(h, collision=_abi.ADE_COLL_CLBM, diffusivity=0.02, geo=ageo)
    a.set_inflow(1.7)
    rng = np.random.default_rng(7)
    g = np.ascontiguousarray((0.1 + 0.05 * rng.random((7, X, Y, Z))).transpose(0, 1, 3, 2)).reshape(-1)
    a.df[0][:] = g
    g0 = a.export_f64(0)
    h.step_coupled(a, output_macro=True)
    out = a.export_f64()
    rho, vx, vy, vz = h.macros()
    for y in range(1, Y - 1):
        for z in range(1, Z - 1):
            if z == 6:
                continue
            u = np.array([vx[3, y, z], vy[3, y, z], vz[3, y, z]])
            cu = C[:7] @ u
            feq = np.array([0.25] + [0.125] * 6) * 1.7 * (1 - 2 * (u @ u) + 4 * cu + 8 * cu * cu)
            assert np.abs(out[:, 3, y, z] - feq).max() < 1e-15
    for x in range(1, X - 1):
        if x == 3:
            continue
        for y in range(1, Y - 1):
            p = np.array([g0[q, x - CX[q], y - CY[q], 6 - CZ[q]] for q in range(7)])
            want = p.copy()
            want[qidx((0, 0, -1))] = p[qidx((0, 0, 1))]
            assert np.array_equal(out[:, x, y, 6], want)


def ade_raw(g):
    """(7, X, Y, Z) populations -> the storage order of one D3Q7 array (y fastest)."""
    return np.ascontiguousarray(g.transpose(0, 1, 3, 2)).reshape(-1)


@pytest.mark.parametrize("coll", [_abi.ADE_COLL_SRT, _abi.ADE_COLL_MRT, _abi.ADE_COLL_CLBM])
def test_d3q7_outflow_right_pulls_the_column_behind(coll):
    """D3Q7 OUTFLOW_RIGHT moves the pull location of every population to x - 1
    (xp = x = xm, d3q7/bc.h:73-74 and 90-92) and collides (bc.h:202).  A
    periodic site pulls c_x = +1 from x - 1, c_x = 0 from x and c_x = -1 from
    x + 1, so a periodic run whose columns x and x + 1 are replaced by the
    column x - 1 must store the same bits at x, with the same NSE velocity
    driving the collision."""
    xo = 5
    geo = np.full((X, Y, Z), GEO.PERIODIC, dtype=np.int16)
    ageo = np.full((X, Y, Z), ADE_GEO.PERIODIC, dtype=np.int16)
    ageo[xo] = ADE_GEO.OUTFLOW_RIGHT
    rng = np.random.default_rng(9)
    g0 = 0.1 + 0.05 * rng.random((7, X, Y, Z))
    gb = g0.copy()
    gb[:, xo] = g0[:, xo - 1]
    gb[:, (xo + 1) % X] = g0[:, xo - 1]
    outs = []
    for amap, g in ((ageo, g0), (None, gb)):
        h, _ = lattice(geo, seed=8)
        a = O.HostAde(h, collision=coll, diffusivity=0.02, geo=amap)
        a.df[0][:] = ade_raw(g)
        h.step_coupled(a, output_macro=True)
        outs.append(a.export_f64())
    assert np.array_equal(outs[0][:, xo], outs[1][:, xo])
    # the plane is not a copy of its input: it collided (the periodic run's
    # uncollided pull of column x - 1 differs from the stored result)
    pull = np.array([[[gb[q, xo - CX[q], (y - CY[q]) % Y, (z - CZ[q]) % Z] for z in range(Z)] for y in range(Y)] for q in range(7)])
    assert np.abs(outs[0][:, xo] - pull).max() > 1e-6


def test_inflow_site_holds_the_profile_equilibrium_for_float_and_double_profiles():
    """An INFLOW site takes rho = 1 and u = inflow(y, z), then f = f_eq
    (bc.h:86-90, NSE_Data_InflowProfile::inflow, lbm_data.h:104-132): after
    one step each inflow site holds the product equilibrium of its profile
    value (the default cumulant leaves it in place).  A float profile (TraitsSP's dreal, inflow_profile_real = F32) is
    read as float: the same values in a double profile give the same bits."""
    geo = np.full((X, Y, Z), GEO.PERIODIC, dtype=np.int16)
    geo[3] = GEO.INFLOW
    rng = np.random.default_rng(9)
    prof32 = [(0.02 * rng.standard_normal((Z, Y))).astype(np.float32) for _ in range(3)]
    out = {}
    for kind in ("f32", "f64"):
        h, _ = lattice(geo, seed=3)
        prof = prof32 if kind == "f32" else [p.astype(np.float64) for p in prof32]
        h.set_inflow(0.5, 0.5, 0.5, profile=prof, profile_real=kind)
        h.step()
        out[kind] = h.export_f64()
    assert np.array_equal(out["f32"], out["f64"])
    f = out["f32"]
    for y in range(Y):
        for z in range(Z):
            u = [float(prof32[k][z, y]) for k in range(3)]
            assert np.abs(f[:, 3, y, z] - O.equilibrium(1.0, *u, kind=0)).max() < 1e-15  # collided: a fixed point
