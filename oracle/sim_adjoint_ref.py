"""Restatement of the per-site operators of the reference's steady SRT adjoint
(sim_adjoint) and of the forward operators they belong to, in numpy FP64.

Test infrastructure only: the tests use it to pin what the reference's
adjoint computes against the exact transposes and derivatives of its forward
step.  Nothing in the product imports it.

Forward (the primary run of sim_adjoint.cu:596-611):
- ``density_velocity``: D3Q27_COMMON::computeDensityAndVelocity without force
  (d3q27/common.h:17-50): rho = sum f, u = j / rho;
- ``feq``: D3Q27_EQ (d3q27/eq.h:13-17 and its eq_* weights): the second-order
  polynomial equilibrium w_q rho (1 - 3/2 u^2 + 3 c.u + 9/2 (c.u)^2);
- ``srt_collide``: D3Q27_SRT::collision without force (d3q27/col_srt.h:13-104):
  f + (feq - f) / tau, tau = 3 nu + 1/2;
- ``bb_inlet``: GEO_INFLOW_BB_LEFT (d3q27/bc.h:143-165): the nine +x
  populations are the site's own unstreamed -x slots minus
  2 * 3 * w * rho * (c_opp . u_in); the forward's rho there is the kernel
  struct's default 1 (defs.h:262; the load of the macro density is commented
  out, bc.h:145).

Adjoint (the adjoint run of sim_adjoint.cu:614-633):
- ``srt_adjoint_collide``: D3Q27_SRT_ADJOINT::collision
  (d3q27/col_srt_adjoint.h:10-109) with D3Q27_EQ_ADJOINT::feq
  (d3q27/eq_adjoint.h:13-27), at the primary run's rho, u;
- ``data_source``: D3Q27_EQ_ADJOINT::collisionWithData (eq_adjoint.h:29-33),
  added at GEO_ADJOINT_FLUID_m sites (bc.h:299-302, common_adjoint.h:124-153);
- ``inlet_gradient_y``: NSE_Data_Adjoint::inflow's gy_profile
  (lbm_data.h:177-184; the x and z components are commented out there);
- ``inlet_extra``: the "extra collision" GEO_ADJOINT_INFLOW_BB_LEFT subtracts
  from all 27 adjoint populations (bc.h:340-376).
"""
from __future__ import annotations

import numpy as np

# D3Q27 in the reference's direction order (defs.h:317-350)
CX = np.array([0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1])
CY = np.array([0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, -1, 1, -1, 1])
CZ = np.array([0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, -1, 1])
C = np.stack([CX, CY, CZ], axis=1).astype(np.float64)
NZ = (CX != 0).astype(int) + (CY != 0) + (CZ != 0)
W = np.array([8 / 27, 2 / 27, 1 / 54, 1 / 216])[NZ]
OPP = np.array([0] + [q + 1 if q % 2 == 1 else q - 1 for q in range(1, 27)])
# the nine directions with c_x = +1, which GEO_INFLOW_BB_LEFT replaces
PLUS_X = np.where(CX == 1)[0]


def density_velocity(f):
    rho = f.sum()
    return rho, (C.T @ f) / rho


def feq(rho, u):
    cu = C @ u
    return W * rho * (1.0 - 1.5 * (u @ u) + 3.0 * cu + 4.5 * cu * cu)


def srt_collide(f, nu):
    tau = 3.0 * nu + 0.5
    rho, u = density_velocity(f)
    return f + (feq(rho, u) - f) / tau


def bb_inlet(own, u_in, rho=1.0):
    """The +x populations of a GEO_INFLOW_BB_LEFT site from its own slots
    `own` (27 values, indexed by direction) and the inflow velocity."""
    out = np.zeros(27)
    for k in PLUS_X:
        # f[k] = own[opp(k)] - 2 w rho 3 (c_opp(k) . u_in), with c_opp(k) = -c_k
        out[k] = own[OPP[k]] - 2.0 * W[k] * rho * 3.0 * (-(C[k] @ u_in))
    return out


def _feq_adjoint(k, q, u):
    """D3Q27_EQ_ADJOINT::feq(c_k, c_q, u) = feq_rho + feq_v."""
    ck, cq = C[k], C[q]
    cku = ck @ u
    feq_rho = 1.0 - 1.5 * (u @ u) + 3.0 * cku + 4.5 * cku * cku
    d = cq - u
    feq_v = -1.5 * 2.0 * (u @ d) + 3.0 * (ck @ d) + 4.5 * 2.0 * cku * (ck @ d)
    return feq_rho + feq_v


def srt_adjoint_collide(fs, rho, u, nu):
    """One adjoint collision: fs_q + (sum_k fs_k w_k feq(k, q, u) - fs_q) / tau."""
    tau = 3.0 * nu + 0.5
    out = np.empty(27)
    for q in range(27):
        loc = sum(fs[k] * W[k] * _feq_adjoint(k, q, u) for k in range(27))
        out[q] = fs[q] + (loc - fs[q]) / tau
    return out


def data_source(rho, u, rho_m, u_m):
    """collisionWithData(c_q, ...) for every q."""
    return np.array([(rho - rho_m) + ((u - u_m) @ (C[q] - u)) / rho for q in range(27)])


def inlet_gradient_y(fs, rho):
    """gy_profile of NSE_Data_Adjoint::inflow at one inlet site."""
    g = 0.0
    for k in PLUS_X:
        if CY[k] != 0:
            g += -CY[k] * fs[k] * 2.0 * W[k] * rho * 3.0
    return g


def inlet_extra(fs_tmp, u):
    """The term GEO_ADJOINT_INFLOW_BB_LEFT subtracts from all 27 adjoint
    populations; fs_tmp are the adjoint's -x values loaded from the site's
    own +x slots, before its collision."""
    r = 0.0
    for k in PLUS_X:
        m = OPP[k]  # the -x direction holding the value loaded from slot k
        r += W[m] * fs_tmp[m] * 2.0 * 3.0 * (C[m] @ u)
    return r
