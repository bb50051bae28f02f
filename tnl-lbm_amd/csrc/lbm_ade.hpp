// Device-side D3Q7 advection-diffusion plugins (TNL-LBM include/lbm3d/d3q7/):
// equilibrium (eq.h:8-54), scalar moment (common.h:16-25), SRT/MRT/CLBM
// collisions (col_srt.h, col_mrt.h, col_clbm.h), A-A/A-B streaming and the
// D3Q7_BC_All switch (bc.h:61-212), plus a source term the reference leaves
// as a TODO (lbm_data.h:229): f_q += w_q S after the collision.
#pragma once

#include "lbm_device.hpp"

namespace lbm {

constexpr double w7(int q)
{
	return q == 0 ? 0.25 : 0.125;
}

struct AdeArgs
{
	void* df0;
	void* df1;
	void* macro;
	const int16_t* map;
	const void* diff;    // per-site diffusivity (macro_real) or null
	const void* source;  // per-site source (macro_real) or null
	double D;            // constant diffusivity
	double inflow_phi;
	double phi_ref, scale;
	int32_t collision;
	int32_t even;
};

enum : int16_t
{
	A_FLUID = 0, A_WALL = 1, A_WALL_BODY = 2, A_SOLID = 3, A_INFLOW = 7, A_OUTFLOW_RIGHT = 8, A_PERIODIC = 9,
	A_NOTHING = 10, A_SYM_TOP = 12, A_SYM_BOTTOM = 13, A_SYM_LEFT = 14, A_SYM_RIGHT = 15, A_SYM_BACK = 16, A_SYM_FRONT = 17
};

template <int ST>
struct Ade7
{
	using S = Storage<ST>;
	using real = typename S::real;
	using elem = typename S::elem;

	// deviation forms store (f_q - w_q phi_ref) * scale; the fixed-point form
	// stores that quantity divided by w_q as a signed-normalised integer.
	// Loads take the inverse scale (isc = 1 / scale), formed once per site.
	static LBM_DEV real ld(const elem* p, int q, real pr, real isc)
	{
		const real v = S::ld(p, (float) (w7(q) / 32767.0));
		if constexpr (S::dev)
			return v * isc + (real) w7(q) * pr;
		else
			return v;
	}
	static LBM_DEV void st(elem* p, int q, real v, real pr, real sc)
	{
		if constexpr (S::dev)
			S::st(p, (v - (real) w7(q) * pr) * sc, (float) (1.0 / w7(q)));
		else
			S::st(p, v);
	}

	// D3Q7_EQ (eq.h:19-53), iCs2 = 4
	template <int q>
	static LBM_DEV real eq(real phi, real vx, real vy, real vz)
	{
		const real cu = (real) CX[q] * vx + (real) CY[q] * vy + (real) CZ[q] * vz;
		const real feq = (real) 1 - (real) 2 * (vx * vx + vy * vy + vz * vz) + (real) 4 * cu + (real) 8 * cu * cu;
		return (real) w7(q) * phi * feq;
	}
	static LBM_DEV void set_eq(real* f, real phi, real vx, real vy, real vz)
	{
		f[0] = eq<0>(phi, vx, vy, vz); f[1] = eq<1>(phi, vx, vy, vz); f[2] = eq<2>(phi, vx, vy, vz);
		f[3] = eq<3>(phi, vx, vy, vz); f[4] = eq<4>(phi, vx, vy, vz); f[5] = eq<5>(phi, vx, vy, vz);
		f[6] = eq<6>(phi, vx, vy, vz);
	}
	// full-way bounce-back of the six moving populations (d3q7/bc.h, GEO_WALL)
	static LBM_DEV void bounce_back(real* f)
	{
#pragma unroll
		for (int q = 1; q < 7; q += 2) {
			const real t = f[q];
			f[q] = f[q + 1];
			f[q + 1] = t;
		}
	}
	static LBM_DEV real density(const real* f)
	{
		real phi = 0;
#pragma unroll
		for (int i = 0; i < 7; i++)
			phi += f[i];
		return phi;
	}

	// mu = C^T lam for the collision f -> C f of a given velocity (phi = sum f
	// is a linear functional of f).  SRT is (1 - omega) I plus the rank-one
	// omega (w E) 1^T; MRT and CLBM are I plus, per axis, two rank-one terms
	// (relaxed moment functional) x (update vector) read off collide below,
	// so the transpose costs about one collision instead of seven.
	static LBM_DEV void collide_adjoint(int coll, const real* lam, real* mu, real vx, real vy, real vz, real D)
	{
		const real omega = (real) 1 / ((real) 0.5 + (real) 4 * D);
		if (coll == TNL_LBM_ADE_COLL_SRT) {
			// f' = (1 - omega) f + omega w_q E_q(u) phi (col_srt.h:13-26)
			const real S = eq<0>((real) 1, vx, vy, vz) * lam[0] + eq<1>((real) 1, vx, vy, vz) * lam[1]
						 + eq<2>((real) 1, vx, vy, vz) * lam[2] + eq<3>((real) 1, vx, vy, vz) * lam[3]
						 + eq<4>((real) 1, vx, vy, vz) * lam[4] + eq<5>((real) 1, vx, vy, vz) * lam[5]
						 + eq<6>((real) 1, vx, vy, vz) * lam[6];
#pragma unroll
			for (int q = 0; q < 7; q++)
				mu[q] = ((real) 1 - omega) * lam[q] + omega * S;
			return;
		}
		// per axis a (velocity v, + and - directions p and m):
		//   first moment  alpha = v 1 + e_m - e_p (relaxed with omega),
		//   second moment beta  = MRT: (v^2 + Cs2) 1 - e_m - e_p
		//                         CLBM: (Cs2 - v^2) 1 + (2v - 1) e_p - (2v + 1) e_m,
		// update vectors b (first) and c = -e_0 + (e_p + e_m)/2 (second), with
		//   b = MRT: (e_p - e_m)/2,  CLBM: -2v e_0 + (v + 1/2) e_p + (v - 1/2) e_m
		// (col_mrt.h:15-50, col_clbm.h:15-84); C^T lam = lam + sum_a
		// [omega (b.lam) alpha + (c.lam) beta].
		const bool clbm = coll != TNL_LBM_ADE_COLL_MRT;
		const real Cs2 = (real) 0.25, h = (real) 0.5;
		const real v[3] = {vx, vy, vz};
		real s1[3], s2[3];
		real common = 0;
#pragma unroll
		for (int a = 0; a < 3; a++) {
			const int p = 1 + 2 * a, m = 2 + 2 * a;
			s2[a] = h * (lam[p] + lam[m]) - lam[0];
			if (clbm) {
				s1[a] = omega * ((v[a] + h) * lam[p] + (v[a] - h) * lam[m] - (real) 2 * v[a] * lam[0]);
				common += s1[a] * v[a] + s2[a] * (Cs2 - v[a] * v[a]);
			}
			else {
				s1[a] = omega * h * (lam[p] - lam[m]);
				common += s1[a] * v[a] + s2[a] * (v[a] * v[a] + Cs2);
			}
		}
		mu[0] = lam[0] + common;
#pragma unroll
		for (int a = 0; a < 3; a++) {
			const int p = 1 + 2 * a, m = 2 + 2 * a;
			if (clbm) {
				mu[p] = lam[p] + common - s1[a] + s2[a] * ((real) 2 * v[a] - (real) 1);
				mu[m] = lam[m] + common + s1[a] - s2[a] * ((real) 2 * v[a] + (real) 1);
			}
			else {
				mu[p] = lam[p] + common - s1[a] - s2[a];
				mu[m] = lam[m] + common + s1[a] - s2[a];
			}
		}
	}

	static LBM_DEV void collide(int coll, real* f, real phi, real vx, real vy, real vz, real D)
	{
		const real omega = (real) 1 / ((real) 0.5 + (real) 4 * D);
		if (coll == TNL_LBM_ADE_COLL_SRT) {
			// col_srt.h:13-26
			real fe[7];
			set_eq(fe, phi, vx, vy, vz);
#pragma unroll
			for (int q = 0; q < 7; q++)
				f[q] += omega * (fe[q] - f[q]);
		}
		else if (coll == TNL_LBM_ADE_COLL_MRT) {
			// col_mrt.h:15-50: first moments relax with omega, second with 1
			const real Cs2 = (real) 0.25;
			real m100 = phi * vx + f[mzz] - f[pzz];
			real m010 = phi * vy + f[zmz] - f[zpz];
			real m001 = phi * vz + f[zzm] - f[zzp];
			const real m200 = phi * (vx * vx + Cs2) - f[mzz] - f[pzz];
			const real m020 = phi * (vy * vy + Cs2) - f[zmz] - f[zpz];
			const real m002 = phi * (vz * vz + Cs2) - f[zzm] - f[zzp];
			m100 *= omega; m010 *= omega; m001 *= omega;
			f[zzz] += -m200 - m020 - m002;
			f[pzz] += (real) 0.5 * (m200 + m100);
			f[zpz] += (real) 0.5 * (m020 + m010);
			f[zzp] += (real) 0.5 * (m002 + m001);
			f[mzz] += (real) 0.5 * (m200 - m100);
			f[zmz] += (real) 0.5 * (m020 - m010);
			f[zzm] += (real) 0.5 * (m002 - m001);
		}
		else {
			// col_clbm.h:15-84: central moments, first order with omega, second with 1
			const real Cs2 = (real) 0.25;
			real k100 = phi * vx + f[mzz] - f[pzz];
			real k010 = phi * vy + f[zmz] - f[zpz];
			real k001 = phi * vz + f[zzm] - f[zzp];
			const real k200 = phi * (Cs2 - vx * vx) + (real) 2 * vx * (f[pzz] - f[mzz]) - f[mzz] - f[pzz];
			const real k020 = phi * (Cs2 - vy * vy) + (real) 2 * vy * (f[zpz] - f[zmz]) - f[zmz] - f[zpz];
			const real k002 = phi * (Cs2 - vz * vz) + (real) 2 * vz * (f[zzp] - f[zzm]) - f[zzm] - f[zzp];
			k100 *= omega; k010 *= omega; k001 *= omega;
			f[zzz] += -(real) 2 * (k100 * vx + k010 * vy + k001 * vz) - k200 - k020 - k002;
			f[pzz] += k100 * vx + (real) 0.5 * (k200 + k100);
			f[zpz] += k010 * vy + (real) 0.5 * (k020 + k010);
			f[zzp] += k001 * vz + (real) 0.5 * (k002 + k001);
			f[mzz] += k100 * vx + (real) 0.5 * (k200 - k100);
			f[zmz] += k010 * vy + (real) 0.5 * (k020 - k010);
			f[zzm] += k001 * vz + (real) 0.5 * (k002 - k001);
		}
	}
};

}  // namespace lbm
