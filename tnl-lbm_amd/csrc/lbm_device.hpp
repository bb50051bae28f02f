// Device-side lattice physics for gfx950: direction tables, storage forms and
// the per-site plugins (density/velocity, equilibria, cumulant collisions).
//
// One site per lane; everything is held in registers.  Reference behaviour:
// TNL-LBM include/lbm3d/d3q27/{common.h, eq_inv_cum.h, eq.h, col_cum.h} and,
// for the deviation storage forms, the well-conditioned family
// d3q27/{common_well.h, eq_well.h, col_cum_well.h}.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/tnl_lbm_amd.h"

#define LBM_DEV __device__ __forceinline__

namespace lbm {

// ---------------------------------------------------------------------------
// Direction numbering of defs.h:317-350.  Indexed with compile-time q only
// (all loops over q are fully unrolled).
enum : int
{
	zzz = 0, pzz = 1, mzz = 2, zpz = 3, zmz = 4, zzp = 5, zzm = 6,
	ppz = 7, mmz = 8, pmz = 9, mpz = 10, pzp = 11, mzm = 12, pzm = 13, mzp = 14,
	zpp = 15, zmm = 16, zpm = 17, zmp = 18,
	ppp = 19, mmm = 20, ppm = 21, mmp = 22, pmp = 23, mpm = 24, pmm = 25, mpp = 26
};
constexpr int CX[27] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1};
constexpr int CY[27] = {0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, -1, 1, -1, 1};
constexpr int CZ[27] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, -1, 1};
constexpr int opp(int q)
{
	return q == 0 ? 0 : ((q & 1) ? q + 1 : q - 1);
}
// index of the direction (a-1, b-1, c-1)
constexpr int qidx(int a, int b, int c)
{
	int r = -1;
	for (int q = 0; q < 27; q++)
		if (CX[q] == a - 1 && CY[q] == b - 1 && CZ[q] == c - 1)
			r = q;
	return r;
}
// 1-D rest weights W(-1)=1/6, W(0)=2/3, W(1)=1/6 and the D3Q27 weights
constexpr double W1[3] = {1.0 / 6.0, 2.0 / 3.0, 1.0 / 6.0};
constexpr double w27(int q)
{
	return W1[CX[q] + 1] * W1[CY[q] + 1] * W1[CZ[q] + 1];
}

// Directions in lexicographic order (x slowest; m < z < p on every axis):
// the summation order of the reference's second-moment and stabiliser sums
// (col_mrt.h:16-29, col_kbc_n.h:233-252).
constexpr int LEX27[27] = {mmm, mmz, mmp, mzm, mzz, mzp, mpm, mpz, mpp, zmm, zmz, zmp, zzm, zzz,
						   zzp, zpm, zpz, zpp, pmm, pmz, pmp, pzm, pzz, pzp, ppm, ppz, ppp};
// Raw-moment sums of the KBC operators as (direction, sign) sequences in the
// reference's summation order (col_kbc_n.h:359-382).
struct MomentTerm
{
	int q, s;
};
constexpr MomentTerm KB_M200[18] = {{mmm, 1}, {mmp, 1}, {mmz, 1}, {mpm, 1}, {mpp, 1}, {mpz, 1}, {mzm, 1}, {mzp, 1}, {mzz, 1},
									{pmm, 1}, {pmp, 1}, {pmz, 1}, {ppm, 1}, {ppp, 1}, {ppz, 1}, {pzm, 1}, {pzp, 1}, {pzz, 1}};
constexpr MomentTerm KB_M020[18] = {{mmm, 1}, {mmp, 1}, {mmz, 1}, {mpm, 1}, {mpp, 1}, {mpz, 1}, {zmm, 1}, {zmp, 1}, {zmz, 1},
									{zpm, 1}, {zpp, 1}, {zpz, 1}, {pmm, 1}, {pmp, 1}, {pmz, 1}, {ppm, 1}, {ppp, 1}, {ppz, 1}};
constexpr MomentTerm KB_M002[18] = {{mmm, 1}, {mmp, 1}, {mzm, 1}, {mzp, 1}, {mpm, 1}, {mpp, 1}, {zmm, 1}, {zmp, 1}, {zzm, 1},
									{zzp, 1}, {zpm, 1}, {zpp, 1}, {pmm, 1}, {pmp, 1}, {pzm, 1}, {pzp, 1}, {ppm, 1}, {ppp, 1}};
constexpr MomentTerm KB_M110[12] = {{mmm, 1},  {mmz, 1},  {mmp, 1},  {mpm, -1}, {mpz, -1}, {mpp, -1},
									{pmm, -1}, {pmz, -1}, {pmp, -1}, {ppm, 1},  {ppz, 1},  {ppp, 1}};
constexpr MomentTerm KB_M101[12] = {{mmm, 1},  {mmp, -1}, {mzm, 1},  {mzp, -1}, {mpm, 1},  {mpp, -1},
									{pmm, -1}, {pmp, 1},  {pzm, -1}, {pzp, 1},  {ppm, -1}, {ppp, 1}};
constexpr MomentTerm KB_M011[12] = {{mmm, 1},  {mmp, -1}, {mpm, -1}, {mpp, 1},  {zmm, 1},  {zmp, -1},
									{zpm, -1}, {zpp, 1},  {pmm, 1},  {pmp, -1}, {ppm, -1}, {ppp, 1}};
constexpr MomentTerm KB_M111[8] = {{ppp, 1}, {mmp, 1}, {mpm, 1}, {mpp, -1}, {pmm, 1}, {pmp, -1}, {ppm, -1}, {mmm, -1}};
constexpr MomentTerm KB_M201[12] = {{ppp, 1},  {mmp, 1}, {mzm, -1}, {mzp, 1}, {mpm, -1}, {mpp, 1},
									{pmm, -1}, {pmp, 1}, {pzm, -1}, {pzp, 1}, {ppm, -1}, {mmm, -1}};
constexpr MomentTerm KB_M102[12] = {{ppp, 1}, {mmp, -1}, {mzm, -1}, {mzp, -1}, {mpm, -1}, {mpp, -1},
									{pmm, 1}, {pmp, 1},  {pzm, 1},  {pzp, 1},  {ppm, 1},  {mmm, -1}};
constexpr MomentTerm KB_M210[12] = {{ppp, 1},  {mmz, -1}, {mmp, -1}, {mpm, 1}, {mpz, 1}, {mpp, 1},
									{pmm, -1}, {pmz, -1}, {pmp, -1}, {ppm, 1}, {ppz, 1}, {mmm, -1}};
constexpr MomentTerm KB_M120[12] = {{ppp, 1}, {mmz, -1}, {mmp, -1}, {mpm, -1}, {mpz, -1}, {mpp, -1},
									{pmm, 1}, {pmz, 1},  {pmp, 1},  {ppm, 1},  {ppz, 1},  {mmm, -1}};
constexpr MomentTerm KB_M021[12] = {{ppp, 1},  {mmp, 1}, {mpm, -1}, {mpp, 1}, {zmm, -1}, {zmp, 1},
									{zpm, -1}, {zpp, 1}, {pmm, -1}, {pmp, 1}, {ppm, -1}, {mmm, -1}};
constexpr MomentTerm KB_M012[12] = {{ppp, 1},  {mmp, -1}, {mpm, 1},  {mpp, 1},  {zmm, -1}, {zmp, -1},
									{zpm, 1},  {zpp, 1},  {pmm, -1}, {pmp, -1}, {ppm, 1},  {mmm, -1}};

// equilibrium used by a collision operator for initialisation and boundaries
// (the LBM_EQ default of each reference plugin): 0 = EQ_INV_CUM, 1 = EQ,
// 2 = EQ_ENTROPIC
constexpr int eq_kind_of(int coll)
{
	return coll <= TNL_LBM_COLL_CUM_2017_ANTIALIAS ? 0
		 : (coll == TNL_LBM_COLL_MRT_LES || coll == TNL_LBM_COLL_KBC_N1 || coll == TNL_LBM_COLL_KBC_C1) ? 1
																											: 2;
}

// ---------------------------------------------------------------------------
// Storage forms.  `real` is the arithmetic type, `elem` the stored type.
template <int ST>
struct Storage;
// nontemporal (streaming) accesses of a stored element through its bit type
template <typename E>
struct Bits;
template <>
struct Bits<double> { using type = uint64_t; };
template <>
struct Bits<float> { using type = uint32_t; };
template <>
struct Bits<_Float16> { using type = uint16_t; };
template <>
struct Bits<__bf16> { using type = uint16_t; };
template <typename E>
LBM_DEV E ld_nt(const E* p)
{
	using B = typename Bits<E>::type;
	return __builtin_bit_cast(E, __builtin_nontemporal_load((const B*) p));
}
template <typename E>
LBM_DEV void st_nt(E* p, E v)
{
	using B = typename Bits<E>::type;
	__builtin_nontemporal_store(__builtin_bit_cast(B, v), (B*) p);
}

template <>
struct Storage<TNL_LBM_STORE_F64>
{
	using elem = double;
	using real = double;
	static constexpr bool dev = false;
	static constexpr bool fixed = false;
	static LBM_DEV real ld(const elem* p, float = 1) { return *p; }
	static LBM_DEV void st(elem* p, real v, float = 1) { *p = v; }
};
template <>
struct Storage<TNL_LBM_STORE_F32>
{
	using elem = float;
	using real = float;
	static constexpr bool dev = false;
	static constexpr bool fixed = false;
	static LBM_DEV real ld(const elem* p, float = 1) { return *p; }
	static LBM_DEV void st(elem* p, real v, float = 1) { *p = v; }
};
// Float elements with double arithmetic: a form of the FP32 adjoint step only
// (its lambda fields and recordings stay float), not a storage form of the
// lattice.
constexpr int STORE_F32_F64MATH = 100;
template <>
struct Storage<STORE_F32_F64MATH>
{
	using elem = float;
	using real = double;
	static constexpr bool dev = false;
	static constexpr bool fixed = false;
	static LBM_DEV real ld(const elem* p, float = 1) { return (double) *p; }
	static LBM_DEV void st(elem* p, real v, float = 1) { *p = (float) v; }
};
template <>
struct Storage<TNL_LBM_STORE_F32_DEV>
{
	using elem = float;
	using real = float;
	static constexpr bool dev = true;
	static constexpr bool fixed = false;
	static LBM_DEV real ld(const elem* p, float = 1) { return *p; }
	static LBM_DEV void st(elem* p, real v, float = 1) { *p = v; }
};
template <>
struct Storage<TNL_LBM_STORE_F16_DEV>
{
	using elem = _Float16;
	using real = float;
	static constexpr bool dev = true;
	static constexpr bool fixed = false;
	static LBM_DEV real ld(const elem* p, float = 1) { return (float) *p; }
	static LBM_DEV void st(elem* p, real v, float = 1) { *p = (_Float16) v; }
};
template <>
struct Storage<TNL_LBM_STORE_BF16_DEV>
{
	using elem = __bf16;
	using real = float;
	static constexpr bool dev = true;
	static constexpr bool fixed = false;
	static LBM_DEV real ld(const elem* p, float = 1) { return (float) *p; }
	static LBM_DEV void st(elem* p, real v, float = 1) { *p = (__bf16) v; }
};
// 16-bit fixed point: raw = round(32767 clamp(g / (w_q R), -1, 1)); the
// caller passes sc = w_q R / 32767 (decode) and inv = 1 / (w_q R) (encode)
template <>
struct Storage<TNL_LBM_STORE_I16_DEV>
{
	using elem = int16_t;
	using real = float;
	static constexpr bool dev = true;
	static constexpr bool fixed = true;
	static LBM_DEV real ld(const elem* p, float sc) { return (float) *p * sc; }
	static LBM_DEV void st(elem* p, real v, float inv) { *p = __builtin_amdgcn_cvt_pknorm_i16(v * inv, 0.0f)[0]; }
	// two values, one conversion
	static LBM_DEV void st2(elem* p0, real v0, float inv0, elem* p1, real v1, float inv1)
	{
		const auto r = __builtin_amdgcn_cvt_pknorm_i16(v0 * inv0, v1 * inv1);
		*p0 = r[0];
		*p1 = r[1];
	}
};
template <>
struct Bits<int16_t> { using type = uint16_t; };
// per-direction scale factors of the fixed-point form (w_q R_n / 32767, 1/(w_q R_n))
// direction class |c_q|^2 (0 rest, 1 faces, 2 edges, 3 corners)
constexpr int qclass(int q)
{
	return (CX[q] != 0) + (CY[q] != 0) + (CZ[q] != 0);
}
// the fixed-point bound R_n per class: rs = R_n / 32767 (decode), invr = 1 / R_n (encode)
struct FixedRange
{
	float rs[4];
	float invr[4];
};
LBM_DEV float fx_sc(int q, const float* rs)
{
	return (float) w27(q) * rs[qclass(q)];
}
LBM_DEV float fx_inv(int q, const float* invr)
{
	return (float) (1.0 / w27(q)) * invr[qclass(q)];
}

// ---------------------------------------------------------------------------
// Absolute form (F64 and F32 storage): the reference plugins.
template <typename real>
struct Abs
{
	// D3Q27_COMMON::computeDensityAndVelocity (common.h:17-50), same grouping
	static LBM_DEV void density_velocity(const real* f, real Fx, real Fy, real Fz, real& rho, real& vx, real& vy, real& vz)
	{
		rho = ((((f[ppp] + f[mmm]) + (f[pmp] + f[mpm])) + ((f[ppm] + f[mmp]) + (f[mpp] + f[pmm])))
			   + (((f[zpp] + f[zmm]) + (f[zpm] + f[zmp])) + ((f[pzp] + f[mzm]) + (f[pzm] + f[mzp]))
				  + ((f[ppz] + f[mmz]) + (f[pmz] + f[mpz])))
			   + ((f[pzz] + f[mzz]) + (f[zpz] + f[zmz]) + (f[zzp] + f[zzm])))
			+ f[zzz];
		const real rinv = (real) 1 / rho;
		vz = ((((f[ppp] - f[mmm]) + (f[mpp] - f[pmm])) + ((f[pmp] - f[mpm]) + (f[mmp] - f[ppm])))
			  + (((f[zpp] - f[zmm]) + (f[zmp] - f[zpm])) + ((f[pzp] - f[mzm]) + (f[mzp] - f[pzm]))) + (f[zzp] - f[zzm])
			  + Fz * (real) 0.5)
			/ rho;
		vx = ((((f[ppp] - f[mmm]) + (f[pmp] - f[mpm])) + ((f[ppm] - f[mmp]) + (f[pmm] - f[mpp])))
			  + (((f[pzp] - f[mzm]) + (f[pzm] - f[mzp])) + ((f[ppz] - f[mmz]) + (f[pmz] - f[mpz]))) + (f[pzz] - f[mzz])
			  + Fx * (real) 0.5)
			/ rho;
		vy = ((((f[ppp] - f[mmm]) + (f[ppm] - f[mmp])) + ((f[mpp] - f[pmm]) + (f[mpm] - f[pmp])))
			  + (((f[ppz] - f[mmz]) + (f[mpz] - f[pmz])) + ((f[zpp] - f[zmm]) + (f[zpm] - f[zmp]))) + (f[zpz] - f[zmz])
			  + Fy * (real) 0.5)
			/ rho;
		(void) rinv;
	}

	// D3Q27_EQ_INV_CUM (eq_inv_cum.h:183-295): rho * W(vx) W(vy) W(vz)
	template <int c>
	static LBM_DEV real invcum_factor(real v)
	{
		if constexpr (c == 0)
			return (real) 3 * v * v - (real) 2;
		else if constexpr (c > 0)
			return (real) 3 * v * v + (real) 3 * v + (real) 1;
		else
			return (real) 3 * v * v - (real) 3 * v + (real) 1;
	}
	template <int q>
	static LBM_DEV real eq_invcum(real rho, real vx, real vy, real vz)
	{
		constexpr int nz = (CX[q] != 0) + (CY[q] != 0) + (CZ[q] != 0);
		constexpr double coef = nz == 0 ? -1.0 / 27.0 : nz == 1 ? 1.0 / 54.0 : nz == 2 ? -1.0 / 108.0 : 1.0 / 216.0;
		return (real) coef * rho * (invcum_factor<CX[q]>(vx) * invcum_factor<CY[q]>(vy) * invcum_factor<CZ[q]>(vz));
	}
	// D3Q27_EQ (eq.h:13-130)
	template <int q>
	static LBM_DEV real eq_poly(real rho, real vx, real vy, real vz)
	{
		constexpr int nz = (CX[q] != 0) + (CY[q] != 0) + (CZ[q] != 0);
		constexpr double w = nz == 0 ? 8.0 / 27.0 : nz == 1 ? 2.0 / 27.0 : nz == 2 ? 1.0 / 54.0 : 1.0 / 216.0;
		const real cu = (real) CX[q] * vx + (real) CY[q] * vy + (real) CZ[q] * vz;
		const real feq = (real) 1 - (real) 1.5 * (vx * vx + vy * vy + vz * vz) + (real) 3 * cu + (real) 4.5 * cu * cu;
		return (real) w * rho * feq;
	}
	// D3Q27_EQ_ENTROPIC (eq_entropic.h:10-216): rho prod_d W(c_d) (2 - s_d) A_d^c_d
	// with s_d = sqrt(1 + 3 u_d^2), A_d = (2 u_d + s_d) / (1 - u_d), evaluated
	// weights first, then the (2 - s_d), then A_d axis by axis
	template <int q>
	static LBM_DEV real eq_entropic(real rho, real vx, real vy, real vz)
	{
		constexpr int c[3] = {CX[q], CY[q], CZ[q]};
		const real v[3] = {vx, vy, vz};
		real r = (real) W1[c[0] + 1] * (real) W1[c[1] + 1] * (real) W1[c[2] + 1];
#pragma unroll
		for (int d = 0; d < 3; d++)
			r = r * ((real) 2 - sqrt((real) 1 + (real) 3 * v[d] * v[d]));
#pragma unroll
		for (int d = 0; d < 3; d++) {
			if (c[d] != 0) {
				const real A = ((real) 2 * v[d] + sqrt((real) 1 + (real) 3 * v[d] * v[d])) / ((real) 1 - v[d]);
				r = c[d] < 0 ? r / A : r * A;
			}
		}
		return rho * r;
	}
	// KIND 0: EQ_INV_CUM (cumulants), 1: EQ (MRT_LES, KBC N1/C1), 2: EQ_ENTROPIC (other KBC)
	template <int KIND, int q>
	static LBM_DEV real eq(real rho, real vx, real vy, real vz)
	{
		if constexpr (KIND == 0)
			return eq_invcum<q>(rho, vx, vy, vz);
		else if constexpr (KIND == 1)
			return eq_poly<q>(rho, vx, vy, vz);
		else
			return eq_entropic<q>(rho, vx, vy, vz);
	}

	// ---- D3Q27_MRT "MRT_LES" (col_mrt.h:7-142) ---------------------------
	// Second moment Pi (sums over the lexicographic direction order), its
	// non-equilibrium part relaxed with the Smagorinsky rate
	//   omega = 2 / (sqrt(tau^2 + 18 C sqrt(Q) / rho) + tau), C = 0.0342,
	//   Q = 2 |Pi_neq|^2, tau = 3 nu + 1/2,
	// and f rebuilt from rho, u and Pi with the 2nd-order Hermite form.
	template <int a, int b>
	static LBM_DEV real lexsum2(const real* f)
	{
		real r = 0;
		bool first = true;
#pragma unroll
		for (int i = 0; i < 27; i++) {
			const int q = LEX27[i];
			const int c[3] = {CX[q], CY[q], CZ[q]};
			const int w = c[a] * c[b];
			if (w != 0) {
				r = first ? (w > 0 ? f[q] : -f[q]) : (w > 0 ? r + f[q] : r - f[q]);
				first = false;
			}
		}
		return r;
	}
	static LBM_DEV void mrt(real* f, real rho, real vx, real vy, real vz, real nu)
	{
		real P[6] = {lexsum2<0, 0>(f), lexsum2<1, 1>(f), lexsum2<2, 2>(f), lexsum2<0, 1>(f), lexsum2<0, 2>(f), lexsum2<1, 2>(f)};
		const real third = (real) (1.0 / 3.0);
		const real Pn[6] = {P[0] - rho * (third + vx * vx), P[1] - rho * (third + vy * vy), P[2] - rho * (third + vz * vz),
							P[3] - rho * vx * vy,           P[4] - rho * vx * vz,           P[5] - rho * vy * vz};
		const real Q = (real) 2 * (Pn[0] * Pn[0] + Pn[1] * Pn[1] + Pn[2] * Pn[2] + (real) 2 * (Pn[3] * Pn[3] + Pn[4] * Pn[4] + Pn[5] * Pn[5]));
		const real tau = (real) 3 * nu + (real) 0.5;
		const real C = (real) 0.0342;
		const real omega = (real) 2 / (sqrt(tau * tau + (real) 2 * C * (real) 3 * (real) 3 * sqrt(Q) / rho) + tau);
#pragma unroll
		for (int k = 0; k < 6; k++)
			P[k] -= omega * Pn[k];
		const real trace = P[0] + P[1] + P[2];
#pragma unroll
		for (int q = 0; q < 27; q++) {
			const real cx = (real) CX[q], cy = (real) CY[q], cz = (real) CZ[q];
			const int nz = (CX[q] != 0) + (CY[q] != 0) + (CZ[q] != 0);
			f[q] = (real) w27(q)
				 * (rho * ((real) 2.5 - (real) 1.5 * (real) nz + (real) 3 * (vx * cx + vy * cy + vz * cz))
					+ (real) 4.5 * (P[0] * (cx * cx) + P[1] * (cy * cy) + P[2] * (cz * cz) + (real) 2 * (P[3] * (cx * cy) + P[4] * (cx * cz) + P[5] * (cy * cz)))
					- (real) 1.5 * trace);
		}
	}

	// ---- KBC N1-N4 / C1-C4 (col_kbc_n.h:254-1272, col_kbc_c.h:283-1301) ---
	// f = k + s + h with the shear part s = D (+T) (+Q), in raw (N) or
	// central (C) moments; beta = 1/(2 nu/cs^2 + 1);
	//   gamma = 1/beta - (2 - 1/beta) <ds|dh> / <dh|dh>,  <x|y> = sum x y / feq;
	//   f -= beta (2 ds + gamma dh) - (1 - beta) S feq,  S = 3 (c - u).F / rho;
	// feq is the product-form equilibrium.  The per-direction quantities are
	// recomputed in the two passes instead of being held (register pressure).
	template <int N>
	static LBM_DEV real msum(const real* f, const MomentTerm (&t)[N])
	{
		real r = t[0].s > 0 ? f[t[0].q] : -f[t[0].q];
#pragma unroll
		for (int i = 1; i < N; i++)
			r = t[i].s > 0 ? r + f[t[i].q] : r - f[t[i].q];
		return r;
	}
	// central forms divide (x / 6), raw forms multiply by the reciprocal
	template <bool DIV>
	static LBM_DEV real scl(real x, real n, real inv)
	{
		if constexpr (DIV)
			return x / n;
		else
			return x * inv;
	}
	template <int q, bool DIV>
	static LBM_DEV real tensD(real Nxz, real Nyz, real Pxy, real Pxz, real Pyz)
	{
		constexpr int cx = CX[q], cy = CY[q], cz = CZ[q];
		constexpr int nz = (cx != 0) + (cy != 0) + (cz != 0);
		const real sixth = (real) (1.0 / 6.0);
		if constexpr (nz == 1) {
			if constexpr (cx != 0)
				return scl<DIV>((real) 2 * Nxz - Nyz, (real) 6, sixth);
			else if constexpr (cy != 0)
				return scl<DIV>(-Nxz + (real) 2 * Nyz, (real) 6, sixth);
			else
				return scl<DIV>(-Nxz - Nyz, (real) 6, sixth);
		}
		else if constexpr (nz == 2) {
			const real P = cz == 0 ? Pxy : cy == 0 ? Pxz : Pyz;
			constexpr int sg = cz == 0 ? cx * cy : cy == 0 ? cx * cz : cy * cz;
			return scl<DIV>(sg > 0 ? P : -P, (real) 4, (real) 0.25);
		}
		else
			return 0;
	}
	template <int q, bool DIV>
	static LBM_DEV real tensT(real T)
	{
		constexpr int nz = (CX[q] != 0) + (CY[q] != 0) + (CZ[q] != 0);
		if constexpr (nz == 0)
			return -T;
		else if constexpr (nz == 1)
			return scl<DIV>(T, (real) 6, (real) (1.0 / 6.0));
		else
			return 0;
	}
	// Qc: xxy xxz xyy yyz xzz yzz xyz
	template <int q, bool DIV>
	static LBM_DEV real tensQ(const real* Qc)
	{
		constexpr int cx = CX[q], cy = CY[q], cz = CZ[q];
		constexpr int nz = (cx != 0) + (cy != 0) + (cz != 0);
		if constexpr (nz == 1) {
			const real a = cx ? Qc[2] + Qc[4] : cy ? Qc[0] + Qc[5] : Qc[1] + Qc[3];
			return scl<DIV>((cx + cy + cz) > 0 ? -a : a, (real) 2, (real) 0.5);
		}
		else if constexpr (nz == 2) {
			constexpr int sa = cz == 0 ? cx : cy == 0 ? cx : cy;
			constexpr int sb = cz == 0 ? cy : cz;
			const real a = cz == 0 ? Qc[2] : cy == 0 ? Qc[4] : Qc[5];
			const real b = cz == 0 ? Qc[0] : cy == 0 ? Qc[1] : Qc[3];
			const real t = sa > 0 ? a : -a;
			return scl<DIV>(sb > 0 ? t + b : t - b, (real) 4, (real) 0.25);
		}
		else if constexpr (nz == 3)
			return scl<DIV>(cx * cy * cz > 0 ? Qc[6] : -Qc[6], (real) 8, (real) 0.125);
		else
			return 0;
	}
	struct KbcParts
	{
		real X[3][3];  // product-equilibrium factors [axis][c + 1]
		real T, Nxz, Nyz, Pxy, Pxz, Pyz, Qc[7];
		real eT, eNxz, eNyz, ePxy, ePxz, ePyz, eQc[7];
	};
	template <int q, bool CENTRAL, bool WT, bool WQ>
	static LBM_DEV void kbc_site(const KbcParts& K, real rho, real fq, real& feq, real& ds, real& dh)
	{
		feq = -rho * K.X[0][CX[q] + 1] * K.X[1][CY[q] + 1] * K.X[2][CZ[q] + 1];
		real s = tensD<q, CENTRAL>(K.Nxz, K.Nyz, K.Pxy, K.Pxz, K.Pyz) - tensD<q, CENTRAL>(K.eNxz, K.eNyz, K.ePxy, K.ePxz, K.ePyz);
		if constexpr (WT)
			s = s + tensT<q, CENTRAL>(K.T) - tensT<q, CENTRAL>(K.eT);
		if constexpr (WQ)
			s = s + tensQ<q, CENTRAL>(K.Qc) - tensQ<q, CENTRAL>(K.eQc);
		ds = s;
		dh = fq - feq - s;
	}
	template <bool CENTRAL, bool WT, bool WQ, int i>
	static LBM_DEV void kbc_gamma_sums(const KbcParts& K, const real* f, real rho, real& num, real& den)
	{
		if constexpr (i < 27) {
			constexpr int q = LEX27[i];
			real feq, ds, dh;
			kbc_site<q, CENTRAL, WT, WQ>(K, rho, f[q], feq, ds, dh);
			const real ifeq = (real) 1 / feq;
			num = i == 0 ? ds * dh * ifeq : num + ds * dh * ifeq;
			den = i == 0 ? dh * dh * ifeq : den + dh * dh * ifeq;
			kbc_gamma_sums<CENTRAL, WT, WQ, i + 1>(K, f, rho, num, den);
		}
	}
	template <bool CENTRAL, bool WT, bool WQ, int q>
	static LBM_DEV void kbc_update(const KbcParts& K, real* f, real rho, const real* v, const real* F, real beta, real gamma)
	{
		if constexpr (q < 27) {
			real feq, ds, dh;
			kbc_site<q, CENTRAL, WT, WQ>(K, rho, f[q], feq, ds, dh);
			constexpr int c[3] = {CX[q], CY[q], CZ[q]};
			real t[3];
#pragma unroll
			for (int d = 0; d < 3; d++)
				t[d] = c[d] == 0 ? -v[d] * F[d] : (c[d] > 0 ? (-v[d] + (real) 1) * F[d] : (-v[d] - (real) 1) * F[d]);
			const real S = ((real) 3 * (t[0] + t[1] + t[2])) / rho;
			f[q] -= beta * ((real) 2 * ds + gamma * dh) - ((real) 1 - beta) * S * feq;
			kbc_update<CENTRAL, WT, WQ, q + 1>(K, f, rho, v, F, beta, gamma);
		}
	}
	template <bool CENTRAL, bool WT, bool WQ>
	static LBM_DEV void kbc(real* f, real rho, real vx, real vy, real vz, real nu, real Fx, real Fy, real Fz)
	{
		const real v[3] = {vx, vy, vz};
		const real F[3] = {Fx, Fy, Fz};
		const real third = (real) (1.0 / 3.0);
		KbcParts K;
#pragma unroll
		for (int d = 0; d < 3; d++) {
			const real z = third - (real) 1 + v[d] * v[d];
			const real p = -(real) 0.5 * (z + (real) 1 + v[d]);
			K.X[d][0] = p + v[d];
			K.X[d][1] = z;
			K.X[d][2] = p;
		}
		const real M200 = msum(f, KB_M200), M020 = msum(f, KB_M020), M002 = msum(f, KB_M002);
		const real M110 = msum(f, KB_M110), M101 = msum(f, KB_M101), M011 = msum(f, KB_M011);
		const real rT = M200 + M020 + M002, rNxz = M200 - M002, rNyz = M020 - M002;
		if constexpr (! CENTRAL) {
			K.T = rT; K.Nxz = rNxz; K.Nyz = rNyz; K.Pxy = M110; K.Pxz = M101; K.Pyz = M011;
			if constexpr (WQ) {
				K.Qc[0] = msum(f, KB_M210); K.Qc[1] = msum(f, KB_M201); K.Qc[2] = msum(f, KB_M120);
				K.Qc[3] = msum(f, KB_M021); K.Qc[4] = msum(f, KB_M102); K.Qc[5] = msum(f, KB_M012);
				K.Qc[6] = msum(f, KB_M111);
			}
			K.eT = rho * ((real) 3 * third + vx * vx + vy * vy + vz * vz);
			K.eNxz = rho * (vx * vx - vz * vz);
			K.eNyz = rho * (vy * vy - vz * vz);
			K.ePxy = rho * vx * vy;
			K.ePxz = rho * vx * vz;
			K.ePyz = rho * vy * vz;
			K.eQc[0] = rho * vy * (third + vx * vx);
			K.eQc[1] = rho * vz * (third + vx * vx);
			K.eQc[2] = rho * vx * (third + vy * vy);
			K.eQc[3] = rho * vz * (third + vy * vy);
			K.eQc[4] = rho * vx * (third + vz * vz);
			K.eQc[5] = rho * vy * (third + vz * vz);
			K.eQc[6] = rho * vx * vy * vz;
		}
		else {
			// central macros of col_kbc_c.h:57-69 as written (the cubic ones
			// carry 3 u^2 and u_x u_y u_z without the factor rho)
			K.T = rT - rho * (vx * vx + vy * vy + vz * vz);
			K.Nxz = rNxz + rho * (vz * vz - vx * vx);
			K.Nyz = rNyz + rho * (vz * vz - vy * vy);
			K.Pxy = M110 - rho * vx * vy;
			K.Pxz = M101 - rho * vx * vz;
			K.Pyz = M011 - rho * vy * vz;
			if constexpr (WQ) {
				const real T = K.T, Nxz = K.Nxz, Nyz = K.Nyz, Pxy = K.Pxy, Pxz = K.Pxz, Pyz = K.Pyz;
				K.Qc[0] = msum(f, KB_M210) - third * ((real) 6 * vx * Pxy + vy * ((real) 3 * vx * vx + (real) 2 * Nxz - Nyz + T));
				K.Qc[1] = msum(f, KB_M201) - third * ((real) 6 * vx * Pxz + vz * ((real) 3 * vx * vx + (real) 2 * Nxz - Nyz + T));
				K.Qc[2] = msum(f, KB_M120) - third * ((real) 6 * vy * Pxy + vx * ((real) 3 * vy * vy + (real) 2 * Nyz - Nxz + T));
				K.Qc[3] = msum(f, KB_M021) - third * ((real) 6 * vy * Pyz + vz * ((real) 3 * vy * vy + (real) 2 * Nyz - Nxz + T));
				K.Qc[4] = msum(f, KB_M102) - third * ((real) 6 * vz * Pxz + vx * ((real) 3 * vz * vz - Nyz - Nxz + T));
				K.Qc[5] = msum(f, KB_M012) - third * ((real) 6 * vz * Pyz + vy * ((real) 3 * vz * vz - Nyz - Nxz + T));
				K.Qc[6] = msum(f, KB_M111) - vx * Pyz - vy * Pxz - vz * Pxy - vx * vy * vz;
			}
			K.eT = rho * (real) 3 * third;
			K.eNxz = K.eNyz = K.ePxy = K.ePxz = K.ePyz = 0;
#pragma unroll
			for (int k = 0; k < 7; k++)
				K.eQc[k] = 0;
		}
		if constexpr (! WQ) {
#pragma unroll
			for (int k = 0; k < 7; k++)
				K.Qc[k] = K.eQc[k] = 0;
		}
		const real beta = (real) 1 / ((real) 2 * nu / third + (real) 1);
		real num = 0, den = 0;
		kbc_gamma_sums<CENTRAL, WT, WQ, 0>(K, f, rho, num, den);
		const real gamma = (real) 1 / beta - ((real) 2 - (real) 1 / beta) * num / den;
		kbc_update<CENTRAL, WT, WQ, 0>(K, f, rho, v, F, beta, gamma);
	}

	// forward 1-D central-moment pass (col_cum.h Eqs 6-14)
	static LBM_DEV void fwd(real fm, real fz, real fp, real v, real v2, real& k0, real& k1, real& k2)
	{
		const real s = fp + fm, d = fp - fm;
		k0 = s + fz;
		k1 = d - v * k0;
		k2 = s - (real) 2 * v * d + v2 * k0;
	}
	// backward 1-D pass (col_cum.h Eqs 88-96)
	static LBM_DEV void bwd(real k0, real k1, real k2, real v, real v2, real& fm, real& fz, real& fp)
	{
		fz = k0 * ((real) 1 - v2) - (real) 2 * v * k1 - k2;
		fm = (k0 * (v2 - v) + k1 * ((real) 2 * v - (real) 1) + k2) * (real) 0.5;
		fp = (k0 * (v2 + v) + k1 * ((real) 2 * v + (real) 1) + k2) * (real) 0.5;
	}

	// D3Q27_CUM::collision (col_cum.h:13-511).  omega1 = 1/(3 nu + 1/2) is
	// precomputed on the host exactly as the reference computes it.
	template <bool G2017, bool AA>
	static LBM_DEV void cumulant(real* f, real rho, real vx, real vy, real vz, real omega1)
	{
		const real rho_inv = (real) 1 / rho;
		const real vx2 = vx * vx, vy2 = vy * vy, vz2 = vz * vz;
		real T[3][3][3], U[3][3][3], k[3][3][3];
#pragma unroll
		for (int a = 0; a < 3; a++)
#pragma unroll
			for (int b = 0; b < 3; b++)
				fwd(f[qidx(a, b, 0)], f[qidx(a, b, 1)], f[qidx(a, b, 2)], vz, vz2, T[a][b][0], T[a][b][1], T[a][b][2]);
#pragma unroll
		for (int a = 0; a < 3; a++)
#pragma unroll
			for (int j = 0; j < 3; j++)
				fwd(T[a][0][j], T[a][1][j], T[a][2][j], vy, vy2, U[a][0][j], U[a][1][j], U[a][2][j]);
#pragma unroll
		for (int i = 0; i < 3; i++)
#pragma unroll
			for (int j = 0; j < 3; j++)
				fwd(U[0][i][j], U[1][i][j], U[2][i][j], vx, vx2, k[0][i][j], k[1][i][j], k[2][i][j]);

		const real one = 1, two = 2, three = 3, four = 4, half = (real) 0.5, third = (real) (1.0 / 3.0);
		const real omega2 = one;
		const real k_000 = k[0][0][0];
		const real C_200 = k[2][0][0], C_020 = k[0][2][0], C_002 = k[0][0][2];
		const real C_110 = k[1][1][0], C_101 = k[1][0][1], C_011 = k[0][1][1];
		real Ks[3][3][3];
#pragma unroll
		for (int a = 0; a < 3; a++)
#pragma unroll
			for (int b = 0; b < 3; b++)
#pragma unroll
				for (int c = 0; c < 3; c++)
					Ks[a][b][c] = 0;
		const real ks_110 = (one - omega1) * C_110;
		const real ks_101 = (one - omega1) * C_101;
		const real ks_011 = (one - omega1) * C_011;
		if constexpr (! G2017) {
			// Eqs 33-35 without antialiasing (col_cum.h:247-259)
			const real Eq33 = (one - omega1) * (C_200 - C_020);
			const real Eq34 = (one - omega1) * (C_200 - C_002);
			const real Eq35 = k_000 * omega2 + (one - omega2) * (C_200 + C_020 + C_002);
			const real ks_200 = third * (Eq33 + Eq34 + Eq35);
			const real ks_020 = third * (-two * Eq33 + Eq34 + Eq35);
			const real ks_002 = third * (Eq33 - two * Eq34 + Eq35);
			// col_cum.h:346-365
			const real ks_211 = (ks_200 * ks_011 + two * ks_101 * ks_110) * rho_inv;
			const real ks_121 = (ks_020 * ks_101 + two * ks_110 * ks_011) * rho_inv;
			const real ks_112 = (ks_002 * ks_110 + two * ks_011 * ks_101) * rho_inv;
			const real ks_220 = (ks_020 * ks_200 + two * ks_110 * ks_110) * rho_inv;
			const real ks_022 = (ks_002 * ks_020 + two * ks_011 * ks_011) * rho_inv;
			const real ks_202 = (ks_200 * ks_002 + two * ks_101 * ks_101) * rho_inv;
			const real ks_222 =
				(ks_200 * ks_022 + ks_020 * ks_202 + ks_002 * ks_220 + four * (ks_011 * ks_211 + ks_101 * ks_121 + ks_110 * ks_112)) * rho_inv
				- ((real) 16 * ks_110 * ks_101 * ks_011 + four * (ks_101 * ks_101 * ks_020 + ks_011 * ks_011 * ks_200 + ks_110 * ks_110 * ks_002)
				   + two * ks_200 * ks_020 * ks_002)
					  * rho_inv * rho_inv;
			Ks[2][0][0] = ks_200; Ks[0][2][0] = ks_020; Ks[0][0][2] = ks_002;
			Ks[2][1][1] = ks_211; Ks[1][2][1] = ks_121; Ks[1][1][2] = ks_112;
			Ks[2][2][0] = ks_220; Ks[0][2][2] = ks_022; Ks[2][0][2] = ks_202;
			Ks[2][2][2] = ks_222;
		}
		else {
			const real k_111 = k[1][1][1];
			const real k_120 = k[1][2][0], k_102 = k[1][0][2], k_210 = k[2][1][0], k_012 = k[0][1][2], k_201 = k[2][0][1], k_021 = k[0][2][1];
			const real k_211 = k[2][1][1], k_121 = k[1][2][1], k_112 = k[1][1][2];
			const real k_220 = k[2][2][0], k_202 = k[2][0][2], k_022 = k[0][2][2];
			const real k_122 = k[1][2][2], k_212 = k[2][1][2], k_221 = k[2][2][1], k_222 = k[2][2][2];
			const real k_200 = C_200, k_020 = C_020, k_002 = C_002, k_110 = C_110, k_101 = C_101, k_011 = C_011;
			// cumulants of order 4-6 (Eqs 51-54 of Geier 2015)
			const real C_211 = k_211 - (k_200 * k_011 + two * k_101 * k_110) * rho_inv;
			const real C_121 = k_121 - (k_020 * k_101 + two * k_110 * k_011) * rho_inv;
			const real C_112 = k_112 - (k_002 * k_110 + two * k_011 * k_101) * rho_inv;
			const real C_220 = k_220 - (k_020 * k_200 + two * k_110 * k_110) * rho_inv;
			const real C_022 = k_022 - (k_002 * k_020 + two * k_011 * k_011) * rho_inv;
			const real C_202 = k_202 - (k_200 * k_002 + two * k_101 * k_101) * rho_inv;
			const real C_122 = k_122 - (k_020 * k_102 + k_002 * k_120 + four * k_011 * k_111 + two * (k_110 * k_012 + k_101 * k_021)) * rho_inv;
			const real C_212 = k_212 - (k_002 * k_210 + k_200 * k_012 + four * k_101 * k_111 + two * (k_011 * k_201 + k_110 * k_102)) * rho_inv;
			const real C_221 = k_221 - (k_200 * k_021 + k_020 * k_201 + four * k_110 * k_111 + two * (k_101 * k_120 + k_011 * k_210)) * rho_inv;
			const real C_222 = k_222
							 - (four * k_111 * k_111 + k_200 * k_022 + k_020 * k_202 + k_002 * k_220
								+ four * (k_011 * k_211 + k_101 * k_121 + k_110 * k_112) + two * (k_120 * k_102 + k_210 * k_012 + k_201 * k_021))
								   * rho_inv
							 + ((real) 16 * k_110 * k_101 * k_011 + four * (k_101 * k_101 * k_020 + k_011 * k_011 * k_200 + k_110 * k_110 * k_002)
								+ two * k_200 * k_020 * k_002)
								   * rho_inv * rho_inv;
			// relaxation rates and limiters of Geier 2017 (col_cum.h:198-227)
			const real o1 = omega1, o2 = omega2;
			const real lambda = (real) 0.01;
			const real omega3 = (real) 8 * (o1 - two) * (o2 * (three * o1 - one) - (real) 5 * o1)
							  / ((real) 8 * ((real) 5 - two * o1) * o1 + o2 * ((real) 8 + o1 * ((real) 9 * o1 - (real) 26)));
			const real omega4 = (real) 8 * (o1 - two) * (o1 + o2 * (three * o1 - (real) 7))
							  / (o2 * ((real) 56 - (real) 42 * o1 + (real) 9 * o1 * o1) - (real) 8 * o1);
			const real omega5 = (real) 24 * (o1 - two)
							  * (four * o1 * o1 + o1 * o2 * ((real) 18 - (real) 13 * o1) + o2 * o2 * (two + o1 * ((real) 6 * o1 - (real) 11)))
							  / ((real) 16 * o1 * o1 * (o1 - (real) 6) - two * o1 * o2 * ((real) 216 + (real) 5 * o1 * ((real) 9 * o1 - (real) 46))
								 + o2 * o2 * (o1 * (three * o1 - (real) 10) * ((real) 15 * o1 - (real) 28) - (real) 48));
			auto lim = [&](real om, real c) { return om + (one - om) * fabs(c) / (rho * lambda + fabs(c)); };
			const real A = (four * o1 * o1 + two * o1 * o2 * (o1 - (real) 6) + o2 * o2 * (o1 * ((real) 10 - three * o1) - four))
						 / (o1 - o2) / (o2 * (two + three * o1) - (real) 8 * o1);
			const real B = (four * o1 * o2 * ((real) 9 * o1 - (real) 16) - four * o1 * o1 - two * o2 * o2 * (two + (real) 9 * o1 * (o1 - two)))
						 * third / (o1 - o2) / (o2 * (two + three * o1) - (real) 8 * o1);
			real Dxu = 0, Dyv = 0, Dzw = 0, DxvDyu = 0, DxwDzu = 0, DywDzv = 0;
			if constexpr (AA) {
				// col_cum.h:237-245, trace term as written (subtracts rho - 1)
				Dxu = -o1 * half * rho_inv * (two * C_200 - C_020 - C_002) - o2 * half * rho_inv * (C_200 + C_020 + C_002 - (-one + rho));
				Dyv = Dxu + (real) 1.5 * o1 * rho_inv * (C_200 - C_020);
				Dzw = Dxu + (real) 1.5 * o1 * rho_inv * (C_200 - C_002);
				DxvDyu = -three * o1 * rho_inv * C_110;
				DxwDzu = -three * o1 * rho_inv * C_101;
				DywDzv = -three * o1 * rho_inv * C_011;
			}
			const real Eq33 = (one - o1) * (C_200 - C_020) - three * rho * (one - o1 * half) * (vx2 * Dxu - vy2 * Dyv);
			const real Eq34 = (one - o1) * (C_200 - C_002) - three * rho * (one - o1 * half) * (vx2 * Dxu - vz2 * Dzw);
			const real Eq35 = k_000 * o2 + (one - o2) * (C_200 + C_020 + C_002) - three * rho * (one - o2 * half) * (vx2 * Dxu + vy2 * Dyv + vz2 * Dzw);
			const real ks_200 = third * (Eq33 + Eq34 + Eq35);
			const real ks_020 = third * (-two * Eq33 + Eq34 + Eq35);
			const real ks_002 = third * (Eq33 - two * Eq34 + Eq35);
			const real Eq117 = (one - lim(omega3, k_120 + k_102)) * (k_120 + k_102);
			const real Eq118 = (one - lim(omega3, k_210 + k_012)) * (k_210 + k_012);
			const real Eq119 = (one - lim(omega3, k_201 + k_021)) * (k_201 + k_021);
			const real Eq120 = (one - lim(omega4, k_120 - k_102)) * (k_120 - k_102);
			const real Eq121 = (one - lim(omega4, k_210 - k_012)) * (k_210 - k_012);
			const real Eq122 = (one - lim(omega4, k_201 - k_021)) * (k_201 - k_021);
			const real ks_120 = half * (Eq120 + Eq117);
			const real ks_102 = half * (-Eq120 + Eq117);
			const real ks_210 = half * (Eq121 + Eq118);
			const real ks_012 = half * (-Eq121 + Eq118);
			const real ks_021 = half * (-Eq122 + Eq119);
			const real ks_201 = half * (Eq122 + Eq119);
			const real ks_111 = (one - lim(omega5, k_111)) * k_111;
			const real visc = (one / o1 - half);
			const real Eq43 = (real) (2.0 / 3.0) * visc * A * rho * (Dxu - two * Dyv + Dzw);
			const real Eq44 = (real) (2.0 / 3.0) * visc * A * rho * (Dxu + Dyv - two * Dzw);
			const real Eq45 = -(real) (4.0 / 3.0) * visc * A * rho * (Dxu + Dyv + Dzw);
			const real Cs_220 = third * (Eq43 + Eq44 + Eq45);
			const real Cs_202 = third * (-Eq43 + Eq45);
			const real Cs_022 = third * (-Eq44 + Eq45);
			const real Cs_211 = -third * visc * B * rho * DywDzv;
			const real Cs_121 = -third * visc * B * rho * DxwDzu;
			const real Cs_112 = -third * visc * B * rho * DxvDyu;
			(void) C_211; (void) C_121; (void) C_112; (void) C_220; (void) C_022; (void) C_202;
			(void) C_122; (void) C_212; (void) C_221; (void) C_222;
			// backward cumulant transform (Eqs 81-84); omega6..10 = 1 zero out the rest
			const real ks_211 = Cs_211 + (ks_200 * ks_011 + two * ks_101 * ks_110) * rho_inv;
			const real ks_121 = Cs_121 + (ks_020 * ks_101 + two * ks_110 * ks_011) * rho_inv;
			const real ks_112 = Cs_112 + (ks_002 * ks_110 + two * ks_011 * ks_101) * rho_inv;
			const real ks_220 = Cs_220 + (ks_020 * ks_200 + two * ks_110 * ks_110) * rho_inv;
			const real ks_022 = Cs_022 + (ks_002 * ks_020 + two * ks_011 * ks_011) * rho_inv;
			const real ks_202 = Cs_202 + (ks_200 * ks_002 + two * ks_101 * ks_101) * rho_inv;
			const real ks_122 = (ks_020 * ks_102 + ks_002 * ks_120 + four * ks_011 * ks_111 + two * (ks_110 * ks_012 + ks_101 * ks_021)) * rho_inv;
			const real ks_212 = (ks_002 * ks_210 + ks_200 * ks_012 + four * ks_101 * ks_111 + two * (ks_011 * ks_201 + ks_110 * ks_102)) * rho_inv;
			const real ks_221 = (ks_200 * ks_021 + ks_020 * ks_201 + four * ks_110 * ks_111 + two * (ks_101 * ks_120 + ks_011 * ks_210)) * rho_inv;
			const real ks_222 =
				(four * ks_111 * ks_111 + ks_200 * ks_022 + ks_020 * ks_202 + ks_002 * ks_220
				 + four * (ks_011 * ks_211 + ks_101 * ks_121 + ks_110 * ks_112) + two * (ks_120 * ks_102 + ks_210 * ks_012 + ks_201 * ks_021))
					* rho_inv
				- ((real) 16 * ks_110 * ks_101 * ks_011 + four * (ks_101 * ks_101 * ks_020 + ks_011 * ks_011 * ks_200 + ks_110 * ks_110 * ks_002)
				   + two * ks_200 * ks_020 * ks_002)
					  * rho_inv * rho_inv;
			Ks[2][0][0] = ks_200; Ks[0][2][0] = ks_020; Ks[0][0][2] = ks_002;
			Ks[1][1][1] = ks_111; Ks[1][2][0] = ks_120; Ks[1][0][2] = ks_102; Ks[2][1][0] = ks_210;
			Ks[0][1][2] = ks_012; Ks[2][0][1] = ks_201; Ks[0][2][1] = ks_021;
			Ks[2][1][1] = ks_211; Ks[1][2][1] = ks_121; Ks[1][1][2] = ks_112;
			Ks[2][2][0] = ks_220; Ks[0][2][2] = ks_022; Ks[2][0][2] = ks_202;
			Ks[1][2][2] = ks_122; Ks[2][1][2] = ks_212; Ks[2][2][1] = ks_221; Ks[2][2][2] = ks_222;
		}
		Ks[1][1][0] = ks_110; Ks[1][0][1] = ks_101; Ks[0][1][1] = ks_011;
		Ks[0][0][0] = k_000;
		Ks[1][0][0] = -k[1][0][0];  // Geier 2017 forcing scheme (col_cum.h:371-373)
		Ks[0][1][0] = -k[0][1][0];
		Ks[0][0][1] = -k[0][0][1];
#pragma unroll
		for (int j = 0; j < 3; j++)
#pragma unroll
			for (int c = 0; c < 3; c++)
				bwd(Ks[0][j][c], Ks[1][j][c], Ks[2][j][c], vx, vx2, U[0][j][c], U[1][j][c], U[2][j][c]);
#pragma unroll
		for (int a = 0; a < 3; a++)
#pragma unroll
			for (int c = 0; c < 3; c++)
				bwd(U[a][0][c], U[a][1][c], U[a][2][c], vy, vy2, T[a][0][c], T[a][1][c], T[a][2][c]);
#pragma unroll
		for (int a = 0; a < 3; a++)
#pragma unroll
			for (int b = 0; b < 3; b++)
				bwd(T[a][b][0], T[a][b][1], T[a][b][2], vz, vz2, f[qidx(a, b, 0)], f[qidx(a, b, 1)], f[qidx(a, b, 2)]);
	}
};

// ---------------------------------------------------------------------------
// Deviation form (g = f - w, float arithmetic): every chain value is carried
// as its deviation from the value it takes at rest (rho = 1, u = 0); rest
// values are products of the 1-D weight moments (1, 0, 1/3).  Same algebra
// as the reference's well-conditioned cumulant (col_cum_well.h) restricted
// to the default relaxation.
struct Dev
{
	using real = float;

	static LBM_DEV void density_velocity(const real* f, real Fx, real Fy, real Fz, real& rho, real& vx, real& vy, real& vz)
	{
		rho = ((((f[ppp] + f[mmm]) + (f[pmp] + f[mpm])) + ((f[ppm] + f[mmp]) + (f[mpp] + f[pmm])))
			   + (((f[zpp] + f[zmm]) + (f[zpm] + f[zmp])) + ((f[pzp] + f[mzm]) + (f[pzm] + f[mzp]))
				  + ((f[ppz] + f[mmz]) + (f[pmz] + f[mpz])))
			   + ((f[pzz] + f[mzz]) + (f[zpz] + f[zmz]) + (f[zzp] + f[zzm])))
			+ f[zzz];
		rho += 1.0f;
		const real rinv = 1.0f / rho;
		vx = ((((f[ppp] - f[mmm]) + (f[pmp] - f[mpm])) + ((f[ppm] - f[mmp]) + (f[pmm] - f[mpp])))
			  + (((f[pzp] - f[mzm]) + (f[pzm] - f[mzp])) + ((f[ppz] - f[mmz]) + (f[pmz] - f[mpz]))) + (f[pzz] - f[mzz])
			  + Fx * 0.5f)
			* rinv;
		vy = ((((f[ppp] - f[mmm]) + (f[ppm] - f[mmp])) + ((f[mpp] - f[pmm]) + (f[mpm] - f[pmp])))
			  + (((f[ppz] - f[mmz]) + (f[mpz] - f[pmz])) + ((f[zpp] - f[zmm]) + (f[zpm] - f[zmp]))) + (f[zpz] - f[zmz])
			  + Fy * 0.5f)
			* rinv;
		vz = ((((f[ppp] - f[mmm]) + (f[mpp] - f[pmm])) + ((f[pmp] - f[mpm]) + (f[mmp] - f[ppm])))
			  + (((f[zpp] - f[zmm]) + (f[zmp] - f[zpm])) + ((f[pzp] - f[mzm]) + (f[mzp] - f[pzm]))) + (f[zzp] - f[zzm])
			  + Fz * 0.5f)
			* rinv;
	}

	template <int c>
	static LBM_DEV real dW(real v)
	{
		if constexpr (c == 0)
			return -v * v;
		else if constexpr (c > 0)
			return 0.5f * (v * v + v);
		else
			return 0.5f * (v * v - v);
	}
	// product-form equilibrium as a deviation: rho*prod(W0+dW) - prod(W0)
	template <int q>
	static LBM_DEV real eq_invcum(real rho, real vx, real vy, real vz)
	{
		constexpr float wx = (float) W1[CX[q] + 1], wy = (float) W1[CY[q] + 1], wz = (float) W1[CZ[q] + 1];
		const real dx = dW<CX[q]>(vx), dy = dW<CY[q]>(vy), dz = dW<CZ[q]>(vz);
		const real d2 = dx * (wy + dy) + wx * dy;
		const real d3 = d2 * (wz + dz) + (wx * wy) * dz;
		return d3 + (rho - 1.0f) * ((wx * wy * wz) + d3);
	}
	template <int q>
	static LBM_DEV real eq_poly(real rho, real vx, real vy, real vz)
	{
		constexpr int nz = (CX[q] != 0) + (CY[q] != 0) + (CZ[q] != 0);
		constexpr float w = nz == 0 ? (float) (8.0 / 27.0) : nz == 1 ? (float) (2.0 / 27.0) : nz == 2 ? (float) (1.0 / 54.0) : (float) (1.0 / 216.0);
		const real cu = (real) CX[q] * vx + (real) CY[q] * vy + (real) CZ[q] * vz;
		const real d = -1.5f * (vx * vx + vy * vy + vz * vz) + 3.0f * cu + 4.5f * cu * cu;
		return w * (d + (rho - 1.0f) * (1.0f + d));
	}
	template <int KIND, int q>
	static LBM_DEV real eq(real rho, real vx, real vy, real vz)
	{
		if constexpr (KIND == 0)
			return eq_invcum<q>(rho, vx, vy, vz);
		else
			return eq_poly<q>(rho, vx, vy, vz);
	}

	static LBM_DEV void fwd(real dm, real dz, real dp, real R0, real v, real v2, real& k0, real& k1, real& k2)
	{
		const real s = dp + dm, d = dp - dm;
		k0 = s + dz;
		const real full = R0 + k0;
		k1 = d - v * full;
		k2 = s - 2.0f * v * d + v2 * full;
	}
	static LBM_DEV void bwd(real k0, real k1, real k2, real R0, real v, real v2, real& dm, real& dz, real& dp)
	{
		const real full = R0 + k0;
		dz = k0 * (1.0f - v2) - 2.0f * v * k1 - k2 - R0 * v2;
		dm = (full * (v2 - v) + k1 * (2.0f * v - 1.0f) + k2) * 0.5f;
		dp = (full * (v2 + v) + k1 * (2.0f * v + 1.0f) + k2) * 0.5f;
	}
	static constexpr float M(int i)
	{
		return i == 0 ? 1.0f : i == 1 ? 0.0f : (float) (1.0 / 3.0);
	}
	static constexpr float Wf(int a)
	{
		return (float) W1[a];
	}

	static LBM_DEV void cumulant(real* g, real rho, real vx, real vy, real vz, real omega1)
	{
		const real drho = rho - 1.0f;
		const real rho_inv = 1.0f / rho;
		const real vx2 = vx * vx, vy2 = vy * vy, vz2 = vz * vz;
		real T[3][3][3], U[3][3][3], k[3][3][3];
#pragma unroll
		for (int a = 0; a < 3; a++)
#pragma unroll
			for (int b = 0; b < 3; b++)
				fwd(g[qidx(a, b, 0)], g[qidx(a, b, 1)], g[qidx(a, b, 2)], Wf(a) * Wf(b), vz, vz2, T[a][b][0], T[a][b][1], T[a][b][2]);
#pragma unroll
		for (int a = 0; a < 3; a++)
#pragma unroll
			for (int j = 0; j < 3; j++)
				fwd(T[a][0][j], T[a][1][j], T[a][2][j], Wf(a) * M(j), vy, vy2, U[a][0][j], U[a][1][j], U[a][2][j]);
#pragma unroll
		for (int i = 0; i < 3; i++)
#pragma unroll
			for (int j = 0; j < 3; j++)
				fwd(U[0][i][j], U[1][i][j], U[2][i][j], M(i) * M(j), vx, vx2, k[0][i][j], k[1][i][j], k[2][i][j]);
		const real third = (float) (1.0 / 3.0), ninth = (float) (1.0 / 9.0);
		const real one_m = 1.0f - omega1;
		const real C_200 = k[2][0][0], C_020 = k[0][2][0], C_002 = k[0][0][2];
		const real Eq33 = one_m * (C_200 - C_020);
		const real Eq34 = one_m * (C_200 - C_002);
		const real Eq35 = drho;  // omega2 = 1
		const real a = third * (Eq33 + Eq34 + Eq35);
		const real b = third * (-2.0f * Eq33 + Eq34 + Eq35);
		const real c = third * (Eq33 - 2.0f * Eq34 + Eq35);
		const real s = one_m * k[1][1][0], p = one_m * k[1][0][1], qq = one_m * k[0][1][1];
		const real ks_211 = ((third + a) * qq + 2.0f * p * s) * rho_inv;
		const real ks_121 = ((third + b) * p + 2.0f * s * qq) * rho_inv;
		const real ks_112 = ((third + c) * s + 2.0f * qq * p) * rho_inv;
		const real ks_220 = ((b * a + 2.0f * s * s) + (b + a) * third - drho * ninth) * rho_inv;
		const real ks_022 = ((c * b + 2.0f * qq * qq) + (c + b) * third - drho * ninth) * rho_inv;
		const real ks_202 = ((a * c + 2.0f * p * p) + (a + c) * third - drho * ninth) * rho_inv;
		const real X1 = (a + b + c) * ninth + (ks_022 + ks_202 + ks_220) * third + a * ks_022 + b * ks_202 + c * ks_220
					  + 4.0f * (qq * ks_211 + p * ks_121 + s * ks_112);
		const real X2 = (float) (2.0 / 3.0) * ((a + b + c) * third + (a * b + b * c + c * a) + 2.0f * (p * p + qq * qq + s * s))
					  + 16.0f * s * p * qq + 4.0f * (p * p * b + qq * qq * a + s * s * c) + 2.0f * a * b * c;
		const real ks_222 = (X1 - (X2 + drho * (drho - 1.0f) * (float) (1.0 / 27.0)) * rho_inv) * rho_inv;
		real Ks[3][3][3];
#pragma unroll
		for (int i = 0; i < 3; i++)
#pragma unroll
			for (int j = 0; j < 3; j++)
#pragma unroll
				for (int l = 0; l < 3; l++)
					Ks[i][j][l] = 0.0f;
		Ks[0][0][0] = drho;
		Ks[1][0][0] = -k[1][0][0];
		Ks[0][1][0] = -k[0][1][0];
		Ks[0][0][1] = -k[0][0][1];
		Ks[2][0][0] = a; Ks[0][2][0] = b; Ks[0][0][2] = c;
		Ks[1][1][0] = s; Ks[1][0][1] = p; Ks[0][1][1] = qq;
		Ks[2][1][1] = ks_211; Ks[1][2][1] = ks_121; Ks[1][1][2] = ks_112;
		Ks[2][2][0] = ks_220; Ks[0][2][2] = ks_022; Ks[2][0][2] = ks_202;
		Ks[2][2][2] = ks_222;
#pragma unroll
		for (int j = 0; j < 3; j++)
#pragma unroll
			for (int l = 0; l < 3; l++)
				bwd(Ks[0][j][l], Ks[1][j][l], Ks[2][j][l], M(j) * M(l), vx, vx2, U[0][j][l], U[1][j][l], U[2][j][l]);
#pragma unroll
		for (int i = 0; i < 3; i++)
#pragma unroll
			for (int l = 0; l < 3; l++)
				bwd(U[i][0][l], U[i][1][l], U[i][2][l], Wf(i) * M(l), vy, vy2, T[i][0][l], T[i][1][l], T[i][2][l]);
#pragma unroll
		for (int i = 0; i < 3; i++)
#pragma unroll
			for (int j = 0; j < 3; j++)
				bwd(T[i][j][0], T[i][j][1], T[i][j][2], Wf(i) * Wf(j), vz, vz2, g[qidx(i, j, 0)], g[qidx(i, j, 1)], g[qidx(i, j, 2)]);
	}
};

// ---------------------------------------------------------------------------
// Lean default cumulant for the FP32-arithmetic forms.
//
// With the default relaxation (col_cum.h:330-366) the post-collision state
// depends only on rho, u, the forcing and the six 2nd-order central moments,
// so the forward transform reduces to the ten raw moments of order <= 2
// (instead of the full 27-moment chain), and the backward x-pass has known
// zeros.  Same mathematics as Abs::cumulant<false,false> / Dev::cumulant.
// DEV = true works on deviations g = f - w (all chain values carried as
// deviations from their rest values); DEV = false on absolute f.
template <bool DEV>
struct FastCum
{
	using real = float;

	static LBM_DEV real rcp(real x)
	{
		return __builtin_amdgcn_rcpf(x);
	}

	// Backward pass along one axis for a column (k0, k1, k2) of central
	// moments with velocity v; R0 is the column's rest moment in the
	// deviation form (a compile-time constant, unused when DEV = false).
	// With  B = v full + k1  and  A = v (B + k1) + k2  (full = R0 + k0):
	//   f_0 = k0 - A,   2 f_- = A - B,   2 f_+ = A + B,
	// six operations per column.  The outer populations are left doubled;
	// after three passes population q carries the factor 2^n(q), n(q) the
	// number of non-zero components of c_q, which unscale(q) removes (the
	// stores fold it into their own conversion or scaling).
	static LBM_DEV void bwd(real k0, real k1, real k2, real R0, real v, real& dm, real& dz, real& dp)
	{
		const real full = DEV ? R0 + k0 : k0;
		const real B = v * full + k1;
		const real A = v * (B + k1) + k2;
		dz = k0 - A;
		dm = A - B;
		dp = A + B;
	}
	// same with k1 = 0
	static LBM_DEV void bwd02(real k0, real k2, real R0, real v, real& dm, real& dz, real& dp)
	{
		const real full = DEV ? R0 + k0 : k0;
		const real B = v * full;
		const real A = v * B + k2;
		dz = k0 - A;
		dm = A - B;
		dp = A + B;
	}
	// k0 = k2 = 0 (and R0 = 0)
	static LBM_DEV void bwd1(real k1, real v, real& dm, real& dz, real& dp)
	{
		const real A = v * (k1 + k1);
		dz = -A;
		dm = A - k1;
		dp = A + k1;
	}
	// k2 = 0 (R0 = 0)
	static LBM_DEV void bwd01(real k0, real k1, real v, real& dm, real& dz, real& dp)
	{
		const real B = v * k0 + k1;
		const real A = v * (B + k1);
		dz = k0 - A;
		dm = A - B;
		dp = A + B;
	}
	// factor 2 carried by the outer positions of a pass, and its inverse
	// over the three passes for population q
	static constexpr float sig(int i)
	{
		return i == 1 ? 1.0f : 2.0f;
	}
	static constexpr float unscale(int q)
	{
		return 1.0f / (sig(CX[q] + 1) * sig(CY[q] + 1) * sig(CZ[q] + 1));
	}

	static constexpr float Wf(int a)
	{
		return (float) W1[a];
	}
	static constexpr float M(int i)
	{
		return i == 0 ? 1.0f : i == 1 ? 0.0f : (float) (1.0 / 3.0);
	}

	// one full step of the default cumulant: f (or g) in, post-collision out;
	// also returns rho, u (pre-collision, force-shifted, as the macros)
	static LBM_DEV void collide(real* f, real Fx, real Fy, real Fz, real omega1, bool do_collide, real& rho, real& vx,
								real& vy, real& vz)
	{
		// z pass sums per (x, y) column: t = f- + f0 + f+, d = f+ - f-, s = f+ + f-
		real t[3][3], d[3][3], sm[3][3];
#pragma unroll
		for (int a = 0; a < 3; a++)
#pragma unroll
			for (int b = 0; b < 3; b++) {
				const real fm = f[qidx(a, b, 0)], fz = f[qidx(a, b, 1)], fp = f[qidx(a, b, 2)];
				sm[a][b] = fp + fm;
				d[a][b] = fp - fm;
				t[a][b] = sm[a][b] + fz;
			}
		// x-collapsed and y-collapsed partial sums
		real tx[3], ty[3];  // tx[a] = sum_b t[a][b]; ty[b] = sum_a t[a][b]
#pragma unroll
		for (int a = 0; a < 3; a++)
			tx[a] = (t[a][0] + t[a][2]) + t[a][1];
#pragma unroll
		for (int b = 0; b < 3; b++)
			ty[b] = (t[0][b] + t[2][b]) + t[1][b];
		const real m0 = (tx[0] + tx[2]) + tx[1];  // sum f (or sum g)
		const real jx = tx[2] - tx[0];
		const real Pxx = tx[2] + tx[0];
		const real jy = ty[2] - ty[0];
		const real Pyy = ty[2] + ty[0];
		real dzs = 0, szs = 0;
#pragma unroll
		for (int a = 0; a < 3; a++)
#pragma unroll
			for (int b = 0; b < 3; b++) {
				dzs += d[a][b];
				szs += sm[a][b];
			}
		const real jz = dzs, Pzz = szs;
		const real Pxy = (t[2][2] + t[0][0]) - (t[2][0] + t[0][2]);
		const real Pxz = ((d[2][0] + d[2][1]) + d[2][2]) - ((d[0][0] + d[0][1]) + d[0][2]);
		const real Pyz = ((d[0][2] + d[1][2]) + d[2][2]) - ((d[0][0] + d[1][0]) + d[2][0]);

		rho = DEV ? m0 + 1.0f : m0;
		const real rinv = rcp(rho);
		vx = (jx + 0.5f * Fx) * rinv;
		vy = (jy + 0.5f * Fy) * rinv;
		vz = (jz + 0.5f * Fz) * rinv;
		if (! do_collide)
			return;
		// 2nd-order central moments: k = P - u j - j u + rho u u  (j = rho u - F/2)
		// (deviation form: the rest value 1/3 of the diagonal is already removed)
		const real k200 = Pxx - vx * (2.0f * jx - rho * vx);
		const real k020 = Pyy - vy * (2.0f * jy - rho * vy);
		const real k002 = Pzz - vz * (2.0f * jz - rho * vz);
		const real k110 = Pxy - vx * jy - vy * (jx - rho * vx);
		const real k101 = Pxz - vx * jz - vz * (jx - rho * vx);
		const real k011 = Pyz - vy * jz - vz * (jy - rho * vy);
		const real third = (float) (1.0 / 3.0), ninth = (float) (1.0 / 9.0);
		const real om = 1.0f - omega1;
		const real Eq33 = om * (k200 - k020);
		const real Eq34 = om * (k200 - k002);
		const real Eq35 = DEV ? m0 : rho;  // k_000 * omega2 (omega2 = 1); DEV: deviation of rho
		const real a = third * (Eq33 + Eq34 + Eq35);
		const real b = third * (-2.0f * Eq33 + Eq34 + Eq35);
		const real c = third * (Eq33 - 2.0f * Eq34 + Eq35);
		const real s = om * k110, p = om * k101, qq = om * k011;
		real ks211, ks121, ks112, ks220, ks022, ks202, ks222;
		if (DEV) {
			const real drho = m0;
			ks211 = ((third + a) * qq + 2.0f * p * s) * rinv;
			ks121 = ((third + b) * p + 2.0f * s * qq) * rinv;
			ks112 = ((third + c) * s + 2.0f * qq * p) * rinv;
			ks220 = ((b * a + 2.0f * s * s) + (b + a) * third - drho * ninth) * rinv;
			ks022 = ((c * b + 2.0f * qq * qq) + (c + b) * third - drho * ninth) * rinv;
			ks202 = ((a * c + 2.0f * p * p) + (a + c) * third - drho * ninth) * rinv;
			const real X1 = (a + b + c) * ninth + (ks022 + ks202 + ks220) * third + a * ks022 + b * ks202 + c * ks220
						  + 4.0f * (qq * ks211 + p * ks121 + s * ks112);
			const real X2 = (float) (2.0 / 3.0) * ((a + b + c) * third + (a * b + b * c + c * a) + 2.0f * (p * p + qq * qq + s * s))
						  + 16.0f * s * p * qq + 4.0f * (p * p * b + qq * qq * a + s * s * c) + 2.0f * a * b * c;
			ks222 = (X1 - (X2 + drho * (drho - 1.0f) * (float) (1.0 / 27.0)) * rinv) * rinv;
		}
		else {
			ks211 = (a * qq + 2.0f * p * s) * rinv;
			ks121 = (b * p + 2.0f * s * qq) * rinv;
			ks112 = (c * s + 2.0f * qq * p) * rinv;
			ks220 = (b * a + 2.0f * s * s) * rinv;
			ks022 = (c * b + 2.0f * qq * qq) * rinv;
			ks202 = (a * c + 2.0f * p * p) * rinv;
			ks222 = (a * ks022 + b * ks202 + c * ks220 + 4.0f * (qq * ks211 + p * ks121 + s * ks112)) * rinv
				  - (16.0f * s * p * qq + 4.0f * (p * p * b + qq * qq * a + s * s * c) + 2.0f * a * b * c) * rinv * rinv;
		}
		const real k000 = DEV ? m0 : rho;
		const real hx = 0.5f * Fx, hy = 0.5f * Fy, hz = 0.5f * Fz;  // ks_100 = -k_100 = F/2
		// x pass: (j, k) columns of Ks[i][j][k]
		real U[3][3][3];
		bwd(k000, hx, a, M(0) * M(0), vx, U[0][0][0], U[1][0][0], U[2][0][0]);
		bwd01(hy, s, vx, U[0][1][0], U[1][1][0], U[2][1][0]);
		bwd02(b, ks220, M(2) * M(0), vx, U[0][2][0], U[1][2][0], U[2][2][0]);
		bwd01(hz, p, vx, U[0][0][1], U[1][0][1], U[2][0][1]);
		bwd02(qq, ks211, 0.0f, vx, U[0][1][1], U[1][1][1], U[2][1][1]);
		bwd1(ks121, vx, U[0][2][1], U[1][2][1], U[2][2][1]);
		bwd02(c, ks202, M(0) * M(2), vx, U[0][0][2], U[1][0][2], U[2][0][2]);
		bwd1(ks112, vx, U[0][1][2], U[1][1][2], U[2][1][2]);
		bwd02(ks022, ks222, M(2) * M(2), vx, U[0][2][2], U[1][2][2], U[2][2][2]);
		// y pass (inputs at x position i carry sig(i))
		real T[3][3][3];
#pragma unroll
		for (int i = 0; i < 3; i++)
#pragma unroll
			for (int l = 0; l < 3; l++)
				bwd(U[i][0][l], U[i][1][l], U[i][2][l], sig(i) * Wf(i) * M(l), vy, T[i][0][l], T[i][1][l], T[i][2][l]);
		// z pass (inputs at (i, j) carry sig(i) sig(j))
#pragma unroll
		for (int i = 0; i < 3; i++)
#pragma unroll
			for (int j = 0; j < 3; j++)
				bwd(T[i][j][0], T[i][j][1], T[i][j][2], sig(i) * sig(j) * Wf(i) * Wf(j), vz, f[qidx(i, j, 0)],
					f[qidx(i, j, 1)], f[qidx(i, j, 2)]);
	}
};

template <int ST>
using Phys = typename std::conditional<Storage<ST>::dev, Dev, Abs<typename Storage<ST>::real>>::type;

}  // namespace lbm
