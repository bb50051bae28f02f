// CPU oracle of the TNL-LBM lattice update -- TEST INFRASTRUCTURE ONLY.
//
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
// this library, as the checker.  The product (libtnl_lbm_amd.so) never calls
// it.  Built with g++ -O2 -ffp-contract=off so that every floating-point
// expression is evaluated exactly as written (no FMA contraction, no
// reassociation), i.e. with the semantics of the reference's host build.
//
// Absolute storage forms (F64, F32) restate the reference per-cell plugins
// with the operation order of the cited lines.  The deviation forms
// (F32_DEV/F16_DEV/BF16_DEV) restate the algebra of the reference's
// well-conditioned family (d3q27/common_well.h, eq_well.h, col_cum_well.h),
// which the reference never wires into a configuration.
//
// Parity pins (see tests/): closed-form equilibrium moments, collision
// invariants and fixed points, 48-element cube-symmetry equivariance, shear
// wave decay rate, A-A(S f0) == A-B(f0), deviation form == absolute form in
// double, D3Q7 diffusion variance growth and advection, MRT_LES Smagorinsky
// rate, the defining relations of the boundary codes INFLOW_LEFT,
// INFLOW_BB_LEFT, OUTFLOW_RIGHT_INTERP, SYM_* and the D3Q7 INFLOW / SYM /
// OUTFLOW_RIGHT codes, the KBC stabiliser as the stationary point of the
// quadratic entropy, the adjoint by dot-product identities, and the three
// cumulant collisions in cumulant space: the cumulants before and after a
// collision, computed independently as derivatives of the log of the discrete
// moment-generating function, obey the reference's relaxation rules (orders
// 2..6, the 2017 limiter with the closed-form rates read back per channel, A
// and B through the fourth-order cumulants).  The reference has no golden
// vectors for this path (its only test fixtures are unrelated IBM matrices
// stored as Git-LFS pointers).

#include "lbm_oracle.h"

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <type_traits>
#include <algorithm>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

thread_local std::string g_err;
int fail(int code, const std::string& msg)
{
	g_err = msg;
	return code;
}

// ---------------------------------------------------------------------------
// Direction table (defs.h:317-350)
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
inline int opp(int q)
{
	return q == 0 ? 0 : ((q & 1) ? q + 1 : q - 1);
}
// index of direction (cx,cy,cz), each in {-1,0,1}
int dir_of(int cx, int cy, int cz)
{
	for (int q = 0; q < 27; q++)
		if (CX[q] == cx && CY[q] == cy && CZ[q] == cz)
			return q;
	return -1;
}

// inflow velocity component k at profile index i = Y z + y: the profile
// array's element (double or float, inflow_profile_real), else the constant
// (NSE_Data_InflowProfile::vel_*, NSE_Data_ConstInflow, lbm_data.h:81-125)
double inflow_at(const tnl_lbm_d3q27_data* SD, int k, int64_t i)
{
	const double c[3] = {SD->inflow_vx, SD->inflow_vy, SD->inflow_vz};
	if (! SD->inflow_profile[k])
		return c[k];
	if (SD->inflow_profile_real == TNL_LBM_REAL_F32)
		return (double) static_cast<const float*>(SD->inflow_profile[k])[i];
	return static_cast<const double*>(SD->inflow_profile[k])[i];
}

// MACRO::computeForcing (macro.h:50-52, kernels.h:103): the per-site force
// field, three planes of macro_real elements, added to the homogeneous force
// that copyQuantities put into the kernel struct (macro.h:80-86)
template <typename real>
void add_site_force(const tnl_lbm_d3q27_data* SD, int mreal, int64_t s, real& fx, real& fy, real& fz)
{
	if (! SD->force_field)
		return;
	if (mreal == TNL_LBM_REAL_F64) {
		const double* F = static_cast<const double*>(SD->force_field);
		fx += (real) F[s];
		fy += (real) F[SD->XYZ + s];
		fz += (real) F[2 * SD->XYZ + s];
	}
	else {
		const float* F = static_cast<const float*>(SD->force_field);
		fx += (real) F[s];
		fy += (real) F[SD->XYZ + s];
		fz += (real) F[2 * SD->XYZ + s];
	}
}

// rest weights w_q = W(cx) W(cy) W(cz), W(0)=2/3, W(+-1)=1/6
double w27(int q)
{
	const double W[3] = {1.0 / 6.0, 2.0 / 3.0, 1.0 / 6.0};
	return W[CX[q] + 1] * W[CY[q] + 1] * W[CZ[q] + 1];
}

// D3Q27 geometry codes (d3q27/bc.h:17-43)
enum : int16_t
{
	GEO_FLUID = 0, GEO_WALL, GEO_INFLOW, GEO_INFLOW_LEFT, GEO_INFLOW_BB_LEFT, GEO_INFLOW_EQ_LEFT, GEO_OUTFLOW_EQ,
	GEO_OUTFLOW_RIGHT, GEO_OUTFLOW_RIGHT_INTERP, GEO_PERIODIC, GEO_NOTHING, GEO_SYM_TOP, GEO_SYM_BOTTOM, GEO_SYM_LEFT,
	GEO_SYM_RIGHT, GEO_SYM_BACK, GEO_SYM_FRONT
};
// D3Q7 geometry codes (d3q7/bc.h:17-36)
enum : int16_t
{
	A_FLUID = 0, A_WALL = 1, A_WALL_BODY = 2, A_SOLID = 3, A_INFLOW = 7, A_OUTFLOW_RIGHT = 8, A_PERIODIC = 9,
	A_NOTHING = 10, A_SYM_TOP = 12, A_SYM_BOTTOM = 13, A_SYM_LEFT = 14, A_SYM_RIGHT = 15, A_SYM_BACK = 16,
	A_SYM_FRONT = 17
};

// ---------------------------------------------------------------------------
// 16-bit formats: round-to-nearest-even conversions
uint16_t f32_to_f16(float v)
{
	uint32_t x;
	std::memcpy(&x, &v, 4);
	const uint32_t sign = (x >> 16) & 0x8000u;
	const uint32_t ax = x & 0x7fffffffu;
	if (ax >= 0x7f800000u)  // inf or nan
		return (uint16_t) (sign | 0x7c00u | (ax > 0x7f800000u ? 0x200u : 0u));
	if (ax >= 0x477ff000u)  // rounds to >= 65520 -> inf
		return (uint16_t) (sign | 0x7c00u);
	if (ax < 0x38800000u) {  // result is subnormal or zero
		if (ax < 0x33000000u)  // < 2^-25: rounds to zero
			return (uint16_t) sign;
		const uint32_t m = (ax & 0x7fffffu) | 0x800000u;
		const int e = (int) (ax >> 23);  // biased exponent, 102..112
		const int shift = 126 - e;       // 14..24
		uint32_t r = m >> shift;
		const uint32_t rem = m & ((1u << shift) - 1u);
		const uint32_t half = 1u << (shift - 1);
		if (rem > half || (rem == half && (r & 1u)))
			r++;
		return (uint16_t) (sign | r);
	}
	uint32_t r = ((ax >> 13) - ((127u - 15u) << 10));
	const uint32_t rem = ax & 0x1fffu;
	if (rem > 0x1000u || (rem == 0x1000u && (r & 1u)))
		r++;
	return (uint16_t) (sign | r);
}
float f16_to_f32(uint16_t h)
{
	const uint32_t sign = (uint32_t) (h & 0x8000u) << 16;
	uint32_t e = (h >> 10) & 0x1fu;
	uint32_t m = h & 0x3ffu;
	uint32_t x;
	if (e == 0) {
		if (m == 0)
			x = sign;
		else {
			e = 127 - 15 + 1;
			while ((m & 0x400u) == 0) {
				m <<= 1;
				e--;
			}
			m &= 0x3ffu;
			x = sign | (e << 23) | (m << 13);
		}
	}
	else if (e == 31)
		x = sign | 0x7f800000u | (m << 13);
	else
		x = sign | ((e + 127 - 15) << 23) | (m << 13);
	float v;
	std::memcpy(&v, &x, 4);
	return v;
}
uint16_t f32_to_bf16(float v)
{
	uint32_t x;
	std::memcpy(&x, &v, 4);
	if ((x & 0x7fffffffu) > 0x7f800000u)
		return (uint16_t) ((x >> 16) | 0x40u);
	const uint32_t lsb = (x >> 16) & 1u;
	x += 0x7fffu + lsb;
	return (uint16_t) (x >> 16);
}
// signed-normalised 16-bit: round(32767 * clamp(x, -1, 1)), ties to even
int16_t f32_to_snorm16(float x)
{
	if (x != x)
		return 0;
	const float c = x < -1.0f ? -1.0f : (x > 1.0f ? 1.0f : x);
	return (int16_t) std::nearbyint(c * 32767.0f);
}
// bound R_n of |f - w_q| / w_q per direction class |c_q|^2 = n for the I16
// deviation form (set per oracle call: store_range_class[n], else store_range)
double g_store_range[4] = {1.0, 1.0, 1.0, 1.0};

float bf16_to_f32(uint16_t h)
{
	const uint32_t x = (uint32_t) h << 16;
	float v;
	std::memcpy(&v, &x, 4);
	return v;
}

// ---------------------------------------------------------------------------
// Lattice indexing (defs.h:90-118, lbm_data.h:45-60)
struct Geometry
{
	int64_t X, Y, Z, ox, oy, oz, XYZ;
	int64_t s(int64_t x, int64_t y, int64_t z) const
	{
		return ((x + ox) * (Z + 2 * oz) + (z + oz)) * (Y + 2 * oy) + (y + oy);
	}
};

// ---------------------------------------------------------------------------
// Per-site register state (defs.h:242-276)
template <typename real>
struct KS27
{
	real f[27];
	real fz = 0, fx = 0, fy = 0;
	real vz = 0, vx = 0, vy = 0;
	real rho = 1, lbmViscosity = 1;
};
template <typename real>
struct KS7
{
	real f[7];
	real vz = 0, vx = 0, vy = 0;
	real phi = 1, lbmViscosity = 1;
	real source = 0;
};

// ===========================================================================
// ABSOLUTE FORM: restatement of the reference plugins
// ===========================================================================
template <typename dreal>
struct Abs
{
	// ciselnik.h:4-60 -- double literals cast to dreal
	static constexpr dreal no1 = (dreal) 1.0, no2 = (dreal) 2.0, no3 = (dreal) 3.0, no4 = (dreal) 4.0,
						   no5 = (dreal) 5.0, no6 = (dreal) 6.0, no7 = (dreal) 7.0, no8 = (dreal) 8.0,
						   no9 = (dreal) 9.0, no10 = (dreal) 10.0, no11 = (dreal) 11.0, no13 = (dreal) 13.0,
						   no15 = (dreal) 15.0, no16 = (dreal) 16.0, no18 = (dreal) 18.0, no24 = (dreal) 24.0,
						   no26 = (dreal) 26.0, no28 = (dreal) 28.0, no42 = (dreal) 42.0, no46 = (dreal) 46.0,
						   no48 = (dreal) 48.0, no56 = (dreal) 56.0, no216 = (dreal) 216.0;
	static constexpr dreal n1o2 = (dreal) 0.5, n3o2 = (dreal) 1.5, n9o2 = (dreal) 4.5, n1o3 = (dreal) (1.0 / 3.0),
						   n2o3 = (dreal) (2.0 / 3.0), n4o3 = (dreal) (4.0 / 3.0), n1o27 = (dreal) (1.0 / 27.0),
						   n1o54 = (dreal) (1.0 / 54.0), n1o108 = (dreal) (1.0 / 108.0),
						   n1o216 = (dreal) (1.0 / 216.0), n2o27 = (dreal) (2.0 / 27.0),
						   n8o27 = (dreal) (8.0 / 27.0), n1o4 = (dreal) 0.25, n1o8 = (dreal) 0.125;

	// D3Q27_COMMON::computeDensityAndVelocity, common.h:17-50 (Geier 2015 App. J order)
	static void densityVelocity(KS27<dreal>& KS)
	{
		const dreal* f = KS.f;
		KS.rho = ((((f[ppp] + f[mmm]) + (f[pmp] + f[mpm])) + ((f[ppm] + f[mmp]) + (f[mpp] + f[pmm])))
				  + (((f[zpp] + f[zmm]) + (f[zpm] + f[zmp])) + ((f[pzp] + f[mzm]) + (f[pzm] + f[mzp]))
					 + ((f[ppz] + f[mmz]) + (f[pmz] + f[mpz])))
				  + ((f[pzz] + f[mzz]) + (f[zpz] + f[zmz]) + (f[zzp] + f[zzm])))
			   + f[zzz];
		KS.vz = ((((f[ppp] - f[mmm]) + (f[mpp] - f[pmm])) + ((f[pmp] - f[mpm]) + (f[mmp] - f[ppm])))
				 + (((f[zpp] - f[zmm]) + (f[zmp] - f[zpm])) + ((f[pzp] - f[mzm]) + (f[mzp] - f[pzm]))) + (f[zzp] - f[zzm])
				 + KS.fz * n1o2)
			  / KS.rho;
		KS.vx = ((((f[ppp] - f[mmm]) + (f[pmp] - f[mpm])) + ((f[ppm] - f[mmp]) + (f[pmm] - f[mpp])))
				 + (((f[pzp] - f[mzm]) + (f[pzm] - f[mzp])) + ((f[ppz] - f[mmz]) + (f[pmz] - f[mpz]))) + (f[pzz] - f[mzz])
				 + KS.fx * n1o2)
			  / KS.rho;
		KS.vy = ((((f[ppp] - f[mmm]) + (f[ppm] - f[mmp])) + ((f[mpp] - f[pmm]) + (f[mpm] - f[pmp])))
				 + (((f[ppz] - f[mmz]) + (f[mpz] - f[pmz])) + ((f[zpp] - f[zmm]) + (f[zpm] - f[zmp]))) + (f[zpz] - f[zmz])
				 + KS.fy * n1o2)
			  / KS.rho;
	}

	// D3Q27_EQ_INV_CUM::eq_*, eq_inv_cum.h:14-135: rho * product of 1-D factors
	static dreal invcum_factor(int c, dreal v)
	{
		if (c == 0)
			return no3 * v * v - no2;
		if (c > 0)
			return no3 * v * v + no3 * v + no1;
		return no3 * v * v - no3 * v + no1;
	}
	static dreal eq_invcum(int q, dreal rho, dreal vx, dreal vy, dreal vz)
	{
		const int nz = (CX[q] != 0) + (CY[q] != 0) + (CZ[q] != 0);
		const dreal coef = nz == 0 ? -n1o27 : nz == 1 ? n1o54 : nz == 2 ? -n1o108 : n1o216;
		return coef * rho * (invcum_factor(CX[q], vx) * invcum_factor(CY[q], vy) * invcum_factor(CZ[q], vz));
	}
	// D3Q27_EQ::feq and eq_*, eq.h:13-130
	static dreal feq(int qx, int qy, int qz, dreal vx, dreal vy, dreal vz)
	{
		return no1 - n3o2 * (vx * vx + vy * vy + vz * vz) + no3 * (qx * vx + qy * vy + qz * vz)
			 + n9o2 * (qx * vx + qy * vy + qz * vz) * (qx * vx + qy * vy + qz * vz);
	}
	static dreal eq_poly(int q, dreal rho, dreal vx, dreal vy, dreal vz)
	{
		const int nz = (CX[q] != 0) + (CY[q] != 0) + (CZ[q] != 0);
		const dreal w = nz == 0 ? n8o27 : nz == 1 ? n2o27 : nz == 2 ? n1o54 : n1o216;
		return w * rho * feq(CX[q], CY[q], CZ[q], vx, vy, vz);
	}
	static dreal eq(int which, int q, dreal rho, dreal vx, dreal vy, dreal vz)
	{
		return which == 0 ? eq_invcum(q, rho, vx, vy, vz) : eq_poly(q, rho, vx, vy, vz);
	}
	// D3Q27_COMMON::setEquilibrium, common.h:61-91
	static void setEquilibrium(KS27<dreal>& KS, int which)
	{
		for (int q = 0; q < 27; q++)
			KS.f[q] = eq(which, q, KS.rho, KS.vx, KS.vy, KS.vz);
	}

	// One 1-D pass of the central-moment chain (col_cum.h:62-167, Eqs 6-14):
	//   k0 = (fp + fm) + fz
	//   k1 = (fp - fm) - v*k0
	//   k2 = (fp + fm) - 2*v*(fp - fm) + v^2*k0
	static void fwd(dreal fm, dreal fz, dreal fp, dreal v, dreal v2, dreal& k0, dreal& k1, dreal& k2)
	{
		k0 = (fp + fm) + fz;
		k1 = (fp - fm) - v * k0;
		k2 = (fp + fm) - no2 * v * (fp - fm) + v2 * k0;
	}
	// One 1-D backward pass (col_cum.h:376-472, Eqs 88-96 of Geier 2015)
	static void bwd(dreal k0, dreal k1, dreal k2, dreal v, dreal v2, dreal& fm, dreal& fz, dreal& fp)
	{
		fz = k0 * (no1 - v2) - no2 * v * k1 - k2;
		fm = (k0 * (v2 - v) + k1 * (no2 * v - no1) + k2) * n1o2;
		fp = (k0 * (v2 + v) + k1 * (no2 * v + no1) + k2) * n1o2;
	}

	// D3Q27_CUM::collision, col_cum.h:13-511.  G2017 = USE_GEIER_CUM_2017,
	// AA = USE_GEIER_CUM_ANTIALIAS (only used together with G2017).
	template <bool G2017, bool AA>
	static void cumulant(KS27<dreal>& KS)
	{
		dreal* f = KS.f;
		const dreal rho = KS.rho;
		const dreal rho_inv = dreal(1.0) / rho;
		const dreal vx = KS.vx, vy = KS.vy, vz = KS.vz;
		const dreal vx_sqr = vx * vx, vy_sqr = vy * vy, vz_sqr = vz * vz;

		// F[a][b][c]: population with c-vector (a-1, b-1, c-1)
		dreal F[3][3][3];
		for (int q = 0; q < 27; q++)
			F[CX[q] + 1][CY[q] + 1][CZ[q] + 1] = f[q];
		// z pass (Eq 6-8), then y pass (Eq 9-11), then x pass (Eq 12-14)
		dreal T[3][3][3], U[3][3][3], k[3][3][3];
		for (int a = 0; a < 3; a++)
			for (int b = 0; b < 3; b++)
				fwd(F[a][b][0], F[a][b][1], F[a][b][2], vz, vz_sqr, T[a][b][0], T[a][b][1], T[a][b][2]);
		for (int a = 0; a < 3; a++)
			for (int j = 0; j < 3; j++)
				fwd(T[a][0][j], T[a][1][j], T[a][2][j], vy, vy_sqr, U[a][0][j], U[a][1][j], U[a][2][j]);
		for (int i = 0; i < 3; i++)
			for (int j = 0; j < 3; j++)
				fwd(U[0][i][j], U[1][i][j], U[2][i][j], vx, vx_sqr, k[0][i][j], k[1][i][j], k[2][i][j]);
		// central moments k_ijk; cumulants of order <= 3 equal them
		const dreal k_000 = k[0][0][0], k_100 = k[1][0][0], k_010 = k[0][1][0], k_001 = k[0][0][1];
		const dreal k_110 = k[1][1][0], k_101 = k[1][0][1], k_011 = k[0][1][1];
		const dreal k_200 = k[2][0][0], k_020 = k[0][2][0], k_002 = k[0][0][2];
		const dreal C_110 = k_110, C_101 = k_101, C_011 = k_011;
		const dreal C_200 = k_200, C_020 = k_020, C_002 = k_002;

		// relaxation definitions (col_cum.h:195-196)
		const dreal omega1 = no1 / (no3 * KS.lbmViscosity + n1o2);
		const dreal omega2 = no1;

		dreal Ks[3][3][3];  // post-collision moments ks_ijk
		const dreal Cs_110 = (no1 - omega1) * C_110;
		const dreal Cs_101 = (no1 - omega1) * C_101;
		const dreal Cs_011 = (no1 - omega1) * C_011;

		if constexpr (G2017) {
			const dreal k_111 = k[1][1][1];
			const dreal k_120 = k[1][2][0], k_102 = k[1][0][2], k_210 = k[2][1][0], k_012 = k[0][1][2];
			const dreal k_201 = k[2][0][1], k_021 = k[0][2][1];
			const dreal k_211 = k[2][1][1], k_121 = k[1][2][1], k_112 = k[1][1][2];
			const dreal k_220 = k[2][2][0], k_202 = k[2][0][2], k_022 = k[0][2][2];
			const dreal k_122 = k[1][2][2], k_212 = k[2][1][2], k_221 = k[2][2][1], k_222 = k[2][2][2];
			const dreal C_111 = k_111, C_120 = k_120, C_102 = k_102, C_210 = k_210, C_012 = k_012, C_201 = k_201,
						C_021 = k_021;
			// Eq 51-54 of Geier 2015 (col_cum.h:169-191)
			const dreal C_211 = k_211 - (k_200 * k_011 + no2 * k_101 * k_110) * rho_inv;
			const dreal C_121 = k_121 - (k_020 * k_101 + no2 * k_110 * k_011) * rho_inv;
			const dreal C_112 = k_112 - (k_002 * k_110 + no2 * k_011 * k_101) * rho_inv;
			const dreal C_220 = k_220 - (k_020 * k_200 + no2 * k_110 * k_110) * rho_inv;
			const dreal C_022 = k_022 - (k_002 * k_020 + no2 * k_011 * k_011) * rho_inv;
			const dreal C_202 = k_202 - (k_200 * k_002 + no2 * k_101 * k_101) * rho_inv;
			const dreal C_122 =
				k_122 - (k_020 * k_102 + k_002 * k_120 + no4 * k_011 * k_111 + no2 * (k_110 * k_012 + k_101 * k_021)) * rho_inv;
			const dreal C_212 =
				k_212 - (k_002 * k_210 + k_200 * k_012 + no4 * k_101 * k_111 + no2 * (k_011 * k_201 + k_110 * k_102)) * rho_inv;
			const dreal C_221 =
				k_221 - (k_200 * k_021 + k_020 * k_201 + no4 * k_110 * k_111 + no2 * (k_101 * k_120 + k_011 * k_210)) * rho_inv;
			const dreal C_222 =
				k_222
				- (no4 * k_111 * k_111 + k_200 * k_022 + k_020 * k_202 + k_002 * k_220
				   + no4 * (k_011 * k_211 + k_101 * k_121 + k_110 * k_112) + no2 * (k_120 * k_102 + k_210 * k_012 + k_201 * k_021))
					  * rho_inv
				+ (no16 * k_110 * k_101 * k_011 + no4 * (k_101 * k_101 * k_020 + k_011 * k_011 * k_200 + k_110 * k_110 * k_002)
				   + no2 * k_200 * k_020 * k_002)
					  * rho_inv * rho_inv;

			// col_cum.h:198-227
			const dreal lambda3 = (dreal) (0.01), lambda4 = (dreal) (0.01), lambda5 = (dreal) (0.01);
			const dreal omega3 = no8 * (omega1 - no2) * (omega2 * (no3 * omega1 - no1) - no5 * omega1)
							   / (no8 * (no5 - no2 * omega1) * omega1 + omega2 * (no8 + omega1 * (no9 * omega1 - no26)));
			const dreal omega120p102 = omega3 + (no1 - omega3) * std::fabs(C_120 + C_102) / (rho * lambda3 + std::fabs(C_120 + C_102));
			const dreal omega210p012 = omega3 + (no1 - omega3) * std::fabs(C_210 + C_012) / (rho * lambda3 + std::fabs(C_210 + C_012));
			const dreal omega201p021 = omega3 + (no1 - omega3) * std::fabs(C_201 + C_021) / (rho * lambda3 + std::fabs(C_201 + C_021));
			const dreal omega4 = no8 * (omega1 - no2) * (omega1 + omega2 * (no3 * omega1 - no7))
							   / (omega2 * (no56 - no42 * omega1 + no9 * omega1 * omega1) - no8 * omega1);
			const dreal omega120m102 = omega4 + (no1 - omega4) * std::fabs(C_120 - C_102) / (rho * lambda4 + std::fabs(C_120 - C_102));
			const dreal omega210m012 = omega4 + (no1 - omega4) * std::fabs(C_210 - C_012) / (rho * lambda4 + std::fabs(C_210 - C_012));
			const dreal omega201m021 = omega4 + (no1 - omega4) * std::fabs(C_201 - C_021) / (rho * lambda4 + std::fabs(C_201 - C_021));
			const dreal omega5 =
				no24 * (omega1 - no2)
				* (no4 * omega1 * omega1 + omega1 * omega2 * (no18 - no13 * omega1) + omega2 * omega2 * (no2 + omega1 * (no6 * omega1 - no11)))
				/ (no16 * omega1 * omega1 * (omega1 - no6) - no2 * omega1 * omega2 * (no216 + no5 * omega1 * (no9 * omega1 - no46))
				   + omega2 * omega2 * (omega1 * (no3 * omega1 - no10) * (no15 * omega1 - no28) - no48));
			const dreal omega111 = omega5 + (no1 - omega5) * std::fabs(C_111) / (rho * lambda5 + std::fabs(C_111));
			const dreal omega6 = no1, omega7 = no1, omega8 = no1, omega9 = no1, omega10 = no1;
			const dreal A = (no4 * omega1 * omega1 + no2 * omega1 * omega2 * (omega1 - no6) + omega2 * omega2 * (omega1 * (no10 - no3 * omega1) - no4))
						  / (omega1 - omega2) / (omega2 * (no2 + no3 * omega1) - no8 * omega1);
			const dreal B =
				(no4 * omega1 * omega2 * (no9 * omega1 - no16) - no4 * omega1 * omega1 - no2 * omega2 * omega2 * (no2 + no9 * omega1 * (omega1 - no2)))
				* n1o3 / (omega1 - omega2) / (omega2 * (no2 + no3 * omega1) - no8 * omega1);

			dreal Dxu = 0, Dyv = 0, Dzw = 0, DxvDyu = 0, DxwDzu = 0, DywDzv = 0;
			if constexpr (AA) {
				// col_cum.h:237-245 -- note the trace term subtracts (rho - 1), as written
				Dxu = -omega1 * n1o2 * rho_inv * (no2 * C_200 - C_020 - C_002)
					- omega2 * n1o2 * rho_inv * (C_200 + C_020 + C_002 - (-no1 + rho));
				Dyv = Dxu + n3o2 * omega1 * rho_inv * (C_200 - C_020);
				Dzw = Dxu + n3o2 * omega1 * rho_inv * (C_200 - C_002);
				DxvDyu = -no3 * omega1 * rho_inv * C_110;
				DxwDzu = -no3 * omega1 * rho_inv * C_101;
				DywDzv = -no3 * omega1 * rho_inv * C_011;
			}
			// Eqs 33-35 (col_cum.h:252-259)
			const dreal Eq33RHS = (no1 - omega1) * (C_200 - C_020) - no3 * rho * (no1 - omega1 * n1o2) * (vx_sqr * Dxu - vy_sqr * Dyv);
			const dreal Eq34RHS = (no1 - omega1) * (C_200 - C_002) - no3 * rho * (no1 - omega1 * n1o2) * (vx_sqr * Dxu - vz_sqr * Dzw);
			const dreal Eq35RHS = k_000 * omega2 + (no1 - omega2) * (C_200 + C_020 + C_002)
								- no3 * rho * (no1 - omega2 * n1o2) * (vx_sqr * Dxu + vy_sqr * Dyv + vz_sqr * Dzw);
			const dreal Cs_200 = n1o3 * (Eq33RHS + Eq34RHS + Eq35RHS);
			const dreal Cs_020 = n1o3 * (-no2 * Eq33RHS + Eq34RHS + Eq35RHS);
			const dreal Cs_002 = n1o3 * (Eq33RHS - no2 * Eq34RHS + Eq35RHS);

			// col_cum.h:262-297
			const dreal Eq117 = (no1 - omega120p102) * (C_120 + C_102);
			const dreal Eq118 = (no1 - omega210p012) * (C_210 + C_012);
			const dreal Eq119 = (no1 - omega201p021) * (C_201 + C_021);
			const dreal Eq120 = (no1 - omega120m102) * (C_120 - C_102);
			const dreal Eq121 = (no1 - omega210m012) * (C_210 - C_012);
			const dreal Eq122 = (no1 - omega201m021) * (C_201 - C_021);
			const dreal Cs_120 = n1o2 * (Eq120 + Eq117);
			const dreal Cs_102 = n1o2 * (-Eq120 + Eq117);
			const dreal Cs_210 = n1o2 * (Eq121 + Eq118);
			const dreal Cs_012 = n1o2 * (-Eq121 + Eq118);
			const dreal Cs_021 = n1o2 * (-Eq122 + Eq119);
			const dreal Cs_201 = n1o2 * (Eq122 + Eq119);
			const dreal Cs_111 = (no1 - omega111) * C_111;
			const dreal Eq43RHS =
				n2o3 * (no1 / omega1 - n1o2) * omega6 * A * rho * (Dxu - no2 * Dyv + Dzw) + (no1 - omega6) * (C_220 - no2 * C_202 + C_022);
			const dreal Eq44RHS =
				n2o3 * (no1 / omega1 - n1o2) * omega6 * A * rho * (Dxu + Dyv - no2 * Dzw) + (no1 - omega6) * (C_220 + C_202 - no2 * C_022);
			const dreal Eq45RHS = -n4o3 * (no1 / omega1 - n1o2) * omega7 * A * rho * (Dxu + Dyv + Dzw) + (no1 - omega7) * (C_220 + C_202 + C_022);
			const dreal Cs_220 = n1o3 * (Eq43RHS + Eq44RHS + Eq45RHS);
			const dreal Cs_202 = n1o3 * (-Eq43RHS + Eq45RHS);
			const dreal Cs_022 = n1o3 * (-Eq44RHS + Eq45RHS);
			const dreal Cs_211 = -n1o3 * (no1 / omega1 - n1o2) * omega8 * B * rho * DywDzv + (no1 - omega8) * C_211;
			const dreal Cs_121 = -n1o3 * (no1 / omega1 - n1o2) * omega8 * B * rho * DxwDzu + (no1 - omega8) * C_121;
			const dreal Cs_112 = -n1o3 * (no1 / omega1 - n1o2) * omega8 * B * rho * DxvDyu + (no1 - omega8) * C_112;
			const dreal Cs_221 = (no1 - omega9) * C_221;
			const dreal Cs_212 = (no1 - omega9) * C_212;
			const dreal Cs_122 = (no1 - omega9) * C_122;
			const dreal Cs_222 = (no1 - omega10) * C_222;

			// backward cumulant transformation, Eqs 81-84 (col_cum.h:303-328); ks = Cs for order <= 3
			const dreal ks_200 = Cs_200, ks_020 = Cs_020, ks_002 = Cs_002, ks_110 = Cs_110, ks_101 = Cs_101, ks_011 = Cs_011;
			const dreal ks_111 = Cs_111, ks_120 = Cs_120, ks_102 = Cs_102, ks_210 = Cs_210, ks_012 = Cs_012, ks_201 = Cs_201,
						ks_021 = Cs_021;
			const dreal ks_211 = Cs_211 + (ks_200 * ks_011 + no2 * ks_101 * ks_110) * rho_inv;
			const dreal ks_121 = Cs_121 + (ks_020 * ks_101 + no2 * ks_110 * ks_011) * rho_inv;
			const dreal ks_112 = Cs_112 + (ks_002 * ks_110 + no2 * ks_011 * ks_101) * rho_inv;
			const dreal ks_220 = Cs_220 + (ks_020 * ks_200 + no2 * ks_110 * ks_110) * rho_inv;
			const dreal ks_022 = Cs_022 + (ks_002 * ks_020 + no2 * ks_011 * ks_011) * rho_inv;
			const dreal ks_202 = Cs_202 + (ks_200 * ks_002 + no2 * ks_101 * ks_101) * rho_inv;
			const dreal ks_122 =
				Cs_122 + (ks_020 * ks_102 + ks_002 * ks_120 + no4 * ks_011 * ks_111 + no2 * (ks_110 * ks_012 + ks_101 * ks_021)) * rho_inv;
			const dreal ks_212 =
				Cs_212 + (ks_002 * ks_210 + ks_200 * ks_012 + no4 * ks_101 * ks_111 + no2 * (ks_011 * ks_201 + ks_110 * ks_102)) * rho_inv;
			const dreal ks_221 =
				Cs_221 + (ks_200 * ks_021 + ks_020 * ks_201 + no4 * ks_110 * ks_111 + no2 * (ks_101 * ks_120 + ks_011 * ks_210)) * rho_inv;
			const dreal ks_222 =
				Cs_222
				+ (no4 * ks_111 * ks_111 + ks_200 * ks_022 + ks_020 * ks_202 + ks_002 * ks_220
				   + no4 * (ks_011 * ks_211 + ks_101 * ks_121 + ks_110 * ks_112) + no2 * (ks_120 * ks_102 + ks_210 * ks_012 + ks_201 * ks_021))
					  * rho_inv
				- (no16 * ks_110 * ks_101 * ks_011 + no4 * (ks_101 * ks_101 * ks_020 + ks_011 * ks_011 * ks_200 + ks_110 * ks_110 * ks_002)
				   + no2 * ks_200 * ks_020 * ks_002)
					  * rho_inv * rho_inv;
			Ks[2][0][0] = ks_200; Ks[0][2][0] = ks_020; Ks[0][0][2] = ks_002;
			Ks[1][1][0] = ks_110; Ks[1][0][1] = ks_101; Ks[0][1][1] = ks_011;
			Ks[1][1][1] = ks_111; Ks[1][2][0] = ks_120; Ks[1][0][2] = ks_102; Ks[2][1][0] = ks_210;
			Ks[0][1][2] = ks_012; Ks[2][0][1] = ks_201; Ks[0][2][1] = ks_021;
			Ks[2][1][1] = ks_211; Ks[1][2][1] = ks_121; Ks[1][1][2] = ks_112;
			Ks[2][2][0] = ks_220; Ks[0][2][2] = ks_022; Ks[2][0][2] = ks_202;
			Ks[1][2][2] = ks_122; Ks[2][1][2] = ks_212; Ks[2][2][1] = ks_221; Ks[2][2][2] = ks_222;
		}
		else {
			// Eqs 33-35 with Dxu = Dyv = Dzw = 0 (col_cum.h:247-259)
			const dreal Dxu = 0, Dyv = 0, Dzw = 0;
			const dreal Eq33RHS = (no1 - omega1) * (C_200 - C_020) - no3 * rho * (no1 - omega1 * n1o2) * (vx_sqr * Dxu - vy_sqr * Dyv);
			const dreal Eq34RHS = (no1 - omega1) * (C_200 - C_002) - no3 * rho * (no1 - omega1 * n1o2) * (vx_sqr * Dxu - vz_sqr * Dzw);
			const dreal Eq35RHS = k_000 * omega2 + (no1 - omega2) * (C_200 + C_020 + C_002)
								- no3 * rho * (no1 - omega2 * n1o2) * (vx_sqr * Dxu + vy_sqr * Dyv + vz_sqr * Dzw);
			const dreal Cs_200 = n1o3 * (Eq33RHS + Eq34RHS + Eq35RHS);
			const dreal Cs_020 = n1o3 * (-no2 * Eq33RHS + Eq34RHS + Eq35RHS);
			const dreal Cs_002 = n1o3 * (Eq33RHS - no2 * Eq34RHS + Eq35RHS);
			const dreal ks_200 = Cs_200, ks_020 = Cs_020, ks_002 = Cs_002, ks_110 = Cs_110, ks_101 = Cs_101, ks_011 = Cs_011;
			// col_cum.h:332-365: 3rd and 5th order to zero, 4th and 6th from products
			const dreal ks_211 = (ks_200 * ks_011 + no2 * ks_101 * ks_110) * rho_inv;
			const dreal ks_121 = (ks_020 * ks_101 + no2 * ks_110 * ks_011) * rho_inv;
			const dreal ks_112 = (ks_002 * ks_110 + no2 * ks_011 * ks_101) * rho_inv;
			const dreal ks_220 = (ks_020 * ks_200 + no2 * ks_110 * ks_110) * rho_inv;
			const dreal ks_022 = (ks_002 * ks_020 + no2 * ks_011 * ks_011) * rho_inv;
			const dreal ks_202 = (ks_200 * ks_002 + no2 * ks_101 * ks_101) * rho_inv;
			const dreal ks_222 =
				(ks_200 * ks_022 + ks_020 * ks_202 + ks_002 * ks_220 + no4 * (ks_011 * ks_211 + ks_101 * ks_121 + ks_110 * ks_112)) * rho_inv
				- (no16 * ks_110 * ks_101 * ks_011 + no4 * (ks_101 * ks_101 * ks_020 + ks_011 * ks_011 * ks_200 + ks_110 * ks_110 * ks_002)
				   + no2 * ks_200 * ks_020 * ks_002)
					  * rho_inv * rho_inv;
			for (int a = 0; a < 3; a++)
				for (int b = 0; b < 3; b++)
					for (int c = 0; c < 3; c++)
						Ks[a][b][c] = 0;
			Ks[2][0][0] = ks_200; Ks[0][2][0] = ks_020; Ks[0][0][2] = ks_002;
			Ks[1][1][0] = ks_110; Ks[1][0][1] = ks_101; Ks[0][1][1] = ks_011;
			Ks[2][1][1] = ks_211; Ks[1][2][1] = ks_121; Ks[1][1][2] = ks_112;
			Ks[2][2][0] = ks_220; Ks[0][2][2] = ks_022; Ks[2][0][2] = ks_202;
			Ks[2][2][2] = ks_222;
		}
		// backward central moment transformation (col_cum.h:369-472)
		Ks[0][0][0] = k_000;
		Ks[1][0][0] = -k_100;  // Geier 2017 forcing scheme
		Ks[0][1][0] = -k_010;
		Ks[0][0][1] = -k_001;
		// x pass (Eq 88-90): Ks[i][j][k] -> U[a][j][k]
		for (int j = 0; j < 3; j++)
			for (int kk = 0; kk < 3; kk++)
				bwd(Ks[0][j][kk], Ks[1][j][kk], Ks[2][j][kk], vx, vx_sqr, U[0][j][kk], U[1][j][kk], U[2][j][kk]);
		// y pass (Eq 91-93)
		for (int a = 0; a < 3; a++)
			for (int kk = 0; kk < 3; kk++)
				bwd(U[a][0][kk], U[a][1][kk], U[a][2][kk], vy, vy_sqr, T[a][0][kk], T[a][1][kk], T[a][2][kk]);
		// z pass (Eq 94-96)
		for (int a = 0; a < 3; a++)
			for (int b = 0; b < 3; b++)
				bwd(T[a][b][0], T[a][b][1], T[a][b][2], vz, vz_sqr, F[a][b][0], F[a][b][1], F[a][b][2]);
		for (int q = 0; q < 27; q++)
			f[q] = F[CX[q] + 1][CY[q] + 1][CZ[q] + 1];
	}

	// ---- D3Q27_EQ_ENTROPIC (eq_entropic.h:10-216) ------------------------
	// rho * prod_d W(c_d) (2 - s_d) A_d^c_d with s_d = sqrt(1 + 3 u_d^2) and
	// A_d = (2 u_d + s_d) / (1 - u_d); the reference multiplies the three
	// weights, then the three (2 - s_d), then divides or multiplies by A_d
	// axis by axis.
	static dreal w1(int c)
	{
		return c == 0 ? n2o3 : (dreal) (1.0 / 6.0);
	}
	static dreal eq_entropic(int q, dreal rho, dreal vx, dreal vy, dreal vz)
	{
		const dreal v[3] = {vx, vy, vz};
		const int c[3] = {CX[q], CY[q], CZ[q]};
		dreal r = w1(c[0]) * w1(c[1]) * w1(c[2]);
		for (int d = 0; d < 3; d++)
			r = r * (no2 - std::sqrt(no1 + no3 * v[d] * v[d]));
		for (int d = 0; d < 3; d++) {
			const dreal A = (no2 * v[d] + std::sqrt(no1 + no3 * v[d] * v[d])) / (no1 - v[d]);
			if (c[d] < 0)
				r = r * no1 / A;
			else if (c[d] > 0)
				r = r * A;
		}
		return rho * r;
	}
	// equilibrium of a collision operator: 0 = EQ_INV_CUM, 1 = EQ, 2 = EQ_ENTROPIC
	static dreal eq_any(int which, int q, dreal rho, dreal vx, dreal vy, dreal vz)
	{
		return which == 2 ? eq_entropic(q, rho, vx, vy, vz) : eq(which, q, rho, vx, vy, vz);
	}
	static void setEquilibriumAny(KS27<dreal>& KS, int which)
	{
		for (int q = 0; q < 27; q++)
			KS.f[q] = eq_any(which, q, KS.rho, KS.vx, KS.vy, KS.vz);
	}

	// ---- moment sums in a fixed direction sequence -------------------------
	struct Term
	{
		int8_t q, s;
	};
	template <size_t N>
	static dreal msum(const dreal* f, const Term (&t)[N])
	{
		dreal r = t[0].s > 0 ? f[t[0].q] : -f[t[0].q];
		for (size_t i = 1; i < N; i++)
			r = t[i].s > 0 ? r + f[t[i].q] : r - f[t[i].q];
		return r;
	}
	// directions in lexicographic order (x slowest; m < z < p on every axis)
	static constexpr int LEX[27] = {mmm, mmz, mmp, mzm, mzz, mzp, mpm, mpz, mpp, zmm, zmz, zmp, zzm, zzz,
									zzp, zpm, zpz, zpp, pmm, pmz, pmp, pzm, pzz, pzp, ppm, ppz, ppp};
	// sum over LEX of c_a c_b f (zero terms skipped): the second-moment sums of col_mrt.h:16-29
	static dreal lexsum2(const dreal* f, int a, int b)
	{
		bool first = true;
		dreal r = 0;
		for (int i = 0; i < 27; i++) {
			const int q = LEX[i];
			const int c[3] = {CX[q], CY[q], CZ[q]};
			const int w = c[a] * c[b];
			if (w == 0)
				continue;
			if (first)
				r = w > 0 ? f[q] : -f[q];
			else
				r = w > 0 ? r + f[q] : r - f[q];
			first = false;
		}
		return r;
	}

	// ---- D3Q27_MRT "MRT_LES" (col_mrt.h:7-142) ---------------------------
	// Only the second moment Pi is relaxed, with a Smagorinsky rate
	//   omega = 2 / (sqrt(tau^2 + 2 C 3 3 sqrt(Q) / rho) + tau),  C = 0.0342,
	// Q = 2 |Pi_neq|^2; f is rebuilt from rho, u and Pi (2nd-order Hermite form).
	static void mrt(KS27<dreal>& KS)
	{
		dreal P[6] = {lexsum2(KS.f, 0, 0), lexsum2(KS.f, 1, 1), lexsum2(KS.f, 2, 2),
					  lexsum2(KS.f, 0, 1), lexsum2(KS.f, 0, 2), lexsum2(KS.f, 1, 2)};  // xx yy zz xy xz yz
		const dreal rho = KS.rho, vx = KS.vx, vy = KS.vy, vz = KS.vz;
		const dreal Pn[6] = {P[0] - rho * (n1o3 + vx * vx), P[1] - rho * (n1o3 + vy * vy), P[2] - rho * (n1o3 + vz * vz),
							 P[3] - rho * vx * vy,          P[4] - rho * vx * vz,          P[5] - rho * vy * vz};
		const dreal Q = no2 * (Pn[0] * Pn[0] + Pn[1] * Pn[1] + Pn[2] * Pn[2] + no2 * (Pn[3] * Pn[3] + Pn[4] * Pn[4] + Pn[5] * Pn[5]));
		const dreal tau = no3 * KS.lbmViscosity + n1o2;
		const dreal C = (dreal) 0.0342;
		const dreal omega = no2 / (std::sqrt(tau * tau + no2 * C * no3 * no3 * std::sqrt(Q) / rho) + tau);
		for (int k = 0; k < 6; k++)
			P[k] -= omega * Pn[k];
		for (int q = 0; q < 27; q++) {
			const dreal cx = (dreal) CX[q], cy = (dreal) CY[q], cz = (dreal) CZ[q];
			const int nz = (CX[q] != 0) + (CY[q] != 0) + (CZ[q] != 0);
			const dreal w = nz == 0 ? n8o27 : nz == 1 ? n2o27 : nz == 2 ? n1o54 : n1o216;
			KS.f[q] = w
					* (rho * ((dreal) 2.5 - n3o2 * (dreal) nz + no3 * (vx * cx + vy * cy + vz * cz))
					   + n9o2 * (P[0] * (cx * cx) + P[1] * (cy * cy) + P[2] * (cz * cz) + no2 * (P[3] * (cx * cy) + P[4] * (cx * cz) + P[5] * (cy * cz)))
					   - n3o2 * (P[0] + P[1] + P[2]));
		}
	}

	// ---- KBC N1-N4 / C1-C4 (col_kbc_n.h:254-1272, col_kbc_c.h:283-1301) ---
	// f = k + s + h; the shear part s holds D (+T) (+Q) in raw (N) or central
	// (C) moments.  beta = 1/(2 nu/cs^2 + 1); the stabiliser
	//   gamma = 1/beta - (2 - 1/beta) <ds|dh> / <dh|dh>,  <x|y> = sum x y / feq;
	// f -= beta (2 ds + gamma dh) - (1 - beta) S feq with the BGK-like force
	// term S = 3 (c - u).F / rho.  feq is the product-form equilibrium.
	// Moment sums follow the reference's direction sequences.
	static constexpr Term M200_[18] = {{mmm, 1}, {mmp, 1}, {mmz, 1}, {mpm, 1}, {mpp, 1}, {mpz, 1}, {mzm, 1}, {mzp, 1}, {mzz, 1},
									   {pmm, 1}, {pmp, 1}, {pmz, 1}, {ppm, 1}, {ppp, 1}, {ppz, 1}, {pzm, 1}, {pzp, 1}, {pzz, 1}};
	static constexpr Term M020_[18] = {{mmm, 1}, {mmp, 1}, {mmz, 1}, {mpm, 1}, {mpp, 1}, {mpz, 1}, {zmm, 1}, {zmp, 1}, {zmz, 1},
									   {zpm, 1}, {zpp, 1}, {zpz, 1}, {pmm, 1}, {pmp, 1}, {pmz, 1}, {ppm, 1}, {ppp, 1}, {ppz, 1}};
	static constexpr Term M002_[18] = {{mmm, 1}, {mmp, 1}, {mzm, 1}, {mzp, 1}, {mpm, 1}, {mpp, 1}, {zmm, 1}, {zmp, 1}, {zzm, 1},
									   {zzp, 1}, {zpm, 1}, {zpp, 1}, {pmm, 1}, {pmp, 1}, {pzm, 1}, {pzp, 1}, {ppm, 1}, {ppp, 1}};
	static constexpr Term M110_[12] = {{mmm, 1},  {mmz, 1},  {mmp, 1},  {mpm, -1}, {mpz, -1}, {mpp, -1},
									   {pmm, -1}, {pmz, -1}, {pmp, -1}, {ppm, 1},  {ppz, 1},  {ppp, 1}};
	static constexpr Term M101_[12] = {{mmm, 1},  {mmp, -1}, {mzm, 1},  {mzp, -1}, {mpm, 1},  {mpp, -1},
									   {pmm, -1}, {pmp, 1},  {pzm, -1}, {pzp, 1},  {ppm, -1}, {ppp, 1}};
	static constexpr Term M011_[12] = {{mmm, 1}, {mmp, -1}, {mpm, -1}, {mpp, 1},  {zmm, 1},  {zmp, -1},
									   {zpm, -1}, {zpp, 1}, {pmm, 1},  {pmp, -1}, {ppm, -1}, {ppp, 1}};
	static constexpr Term M111_[8] = {{ppp, 1}, {mmp, 1}, {mpm, 1}, {mpp, -1}, {pmm, 1}, {pmp, -1}, {ppm, -1}, {mmm, -1}};
	static constexpr Term M201_[12] = {{ppp, 1},  {mmp, 1}, {mzm, -1}, {mzp, 1}, {mpm, -1}, {mpp, 1},
									   {pmm, -1}, {pmp, 1}, {pzm, -1}, {pzp, 1}, {ppm, -1}, {mmm, -1}};
	static constexpr Term M102_[12] = {{ppp, 1}, {mmp, -1}, {mzm, -1}, {mzp, -1}, {mpm, -1}, {mpp, -1},
									   {pmm, 1}, {pmp, 1},  {pzm, 1},  {pzp, 1},  {ppm, 1},  {mmm, -1}};
	static constexpr Term M210_[12] = {{ppp, 1},  {mmz, -1}, {mmp, -1}, {mpm, 1}, {mpz, 1}, {mpp, 1},
									   {pmm, -1}, {pmz, -1}, {pmp, -1}, {ppm, 1}, {ppz, 1}, {mmm, -1}};
	static constexpr Term M120_[12] = {{ppp, 1}, {mmz, -1}, {mmp, -1}, {mpm, -1}, {mpz, -1}, {mpp, -1},
									   {pmm, 1}, {pmz, 1},  {pmp, 1},  {ppm, 1},  {ppz, 1},  {mmm, -1}};
	static constexpr Term M021_[12] = {{ppp, 1},  {mmp, 1},  {mpm, -1}, {mpp, 1}, {zmm, -1}, {zmp, 1},
									   {zpm, -1}, {zpp, 1},  {pmm, -1}, {pmp, 1}, {ppm, -1}, {mmm, -1}};
	static constexpr Term M012_[12] = {{ppp, 1},  {mmp, -1}, {mpm, 1},  {mpp, 1}, {zmm, -1}, {zmp, -1},
									   {zpm, 1},  {zpp, 1},  {pmm, -1}, {pmp, -1}, {ppm, 1}, {mmm, -1}};

	// value of one of the D, T, Q tensors at direction q; `div` selects the
	// central-moment macros' divisions (x / 6) over the raw ones' products (x * 1/6)
	static dreal scale(dreal x, dreal n, dreal inv, bool div)
	{
		return div ? x / n : x * inv;
	}
	static dreal tensD(int q, dreal Nxz, dreal Nyz, dreal Pxy, dreal Pxz, dreal Pyz, bool div)
	{
		const int cx = CX[q], cy = CY[q], cz = CZ[q];
		const int nz = (cx != 0) + (cy != 0) + (cz != 0);
		if (nz == 1) {
			if (cx)
				return scale(no2 * Nxz - Nyz, no6, (dreal) (1.0 / 6.0), div);
			if (cy)
				return scale(-Nxz + no2 * Nyz, no6, (dreal) (1.0 / 6.0), div);
			return scale(-Nxz - Nyz, no6, (dreal) (1.0 / 6.0), div);
		}
		if (nz == 2) {
			const dreal P = cz == 0 ? Pxy : cy == 0 ? Pxz : Pyz;
			const int s = cz == 0 ? cx * cy : cy == 0 ? cx * cz : cy * cz;
			return scale(s > 0 ? P : -P, no4, n1o4, div);
		}
		return 0;
	}
	static dreal tensT(int q, dreal T, bool div)
	{
		const int nz = (CX[q] != 0) + (CY[q] != 0) + (CZ[q] != 0);
		if (nz == 0)
			return -T;
		return nz == 1 ? scale(T, no6, (dreal) (1.0 / 6.0), div) : (dreal) 0;
	}
	// Q components: xxy xxz xyy yyz xzz yzz xyz
	static dreal tensQ(int q, const dreal* Qc, bool div)
	{
		const int cx = CX[q], cy = CY[q], cz = CZ[q];
		const int nz = (cx != 0) + (cy != 0) + (cz != 0);
		enum { xxy, xxz, xyy, yyz, xzz, yzz, xyz };
		if (nz == 1) {
			const dreal a = cx ? Qc[xyy] + Qc[xzz] : cy ? Qc[xxy] + Qc[yzz] : Qc[xxz] + Qc[yyz];
			const int c = cx + cy + cz;
			return scale(c > 0 ? -a : a, no2, n1o2, div);
		}
		if (nz == 2) {
			dreal a, b;
			int sa, sb;
			if (cz == 0) {
				a = Qc[xyy]; sa = cx; b = Qc[xxy]; sb = cy;
			}
			else if (cy == 0) {
				a = Qc[xzz]; sa = cx; b = Qc[xxz]; sb = cz;
			}
			else {
				a = Qc[yzz]; sa = cy; b = Qc[yyz]; sb = cz;
			}
			const dreal t = (sa > 0 ? a : -a);
			return scale(sb > 0 ? t + b : t - b, no4, n1o4, div);
		}
		if (nz == 3)
			return scale(cx * cy * cz > 0 ? Qc[xyz] : -Qc[xyz], no8, n1o8, div);
		return 0;
	}

	template <bool CENTRAL, bool WT, bool WQ>
	static void kbc(KS27<dreal>& KS)
	{
		const dreal rho = KS.rho, vx = KS.vx, vy = KS.vy, vz = KS.vz;
		const dreal* f = KS.f;
		// product-form equilibrium (col_kbc_n.h:289-330)
		const dreal v[3] = {vx, vy, vz};
		dreal Xs[3][3];  // [axis][c + 1]
		for (int d = 0; d < 3; d++) {
			const dreal z = n1o3 - no1 + v[d] * v[d];
			const dreal p = -n1o2 * (z + no1 + v[d]);
			Xs[d][0] = p + v[d];
			Xs[d][1] = z;
			Xs[d][2] = p;
		}
		dreal feq[27], ifeq[27];
		for (int q = 0; q < 27; q++) {
			feq[q] = -rho * Xs[0][CX[q] + 1] * Xs[1][CY[q] + 1] * Xs[2][CZ[q] + 1];
			ifeq[q] = no1 / feq[q];
		}
		const dreal M200 = msum(f, M200_), M020 = msum(f, M020_), M002 = msum(f, M002_);
		const dreal M110 = msum(f, M110_), M101 = msum(f, M101_), M011 = msum(f, M011_);
		const dreal M111 = msum(f, M111_), M201 = msum(f, M201_), M102 = msum(f, M102_), M210 = msum(f, M210_);
		const dreal M120 = msum(f, M120_), M021 = msum(f, M021_), M012 = msum(f, M012_);
		const dreal rT = M200 + M020 + M002, rNxz = M200 - M002, rNyz = M020 - M002;
		// tensor arguments of s and of its equilibrium
		dreal T, Nxz, Nyz, Pxy, Pxz, Pyz, Qc[7];
		dreal eT, eNxz, eNyz, ePxy, ePxz, ePyz, eQc[7];
		if constexpr (! CENTRAL) {
			T = rT; Nxz = rNxz; Nyz = rNyz; Pxy = M110; Pxz = M101; Pyz = M011;
			const dreal q7[7] = {M210, M201, M120, M021, M102, M012, M111};
			for (int k = 0; k < 7; k++)
				Qc[k] = q7[k];
			eT = rho * (no3 * n1o3 + vx * vx + vy * vy + vz * vz);
			eNxz = rho * (vx * vx - vz * vz);
			eNyz = rho * (vy * vy - vz * vz);
			ePxy = rho * vx * vy;
			ePxz = rho * vx * vz;
			ePyz = rho * vy * vz;
			eQc[0] = rho * vy * (n1o3 + vx * vx);
			eQc[1] = rho * vz * (n1o3 + vx * vx);
			eQc[2] = rho * vx * (n1o3 + vy * vy);
			eQc[3] = rho * vz * (n1o3 + vy * vy);
			eQc[4] = rho * vx * (n1o3 + vz * vz);
			eQc[5] = rho * vy * (n1o3 + vz * vz);
			eQc[6] = rho * vx * vy * vz;
		}
		else {
			// central macros of col_kbc_c.h:57-69, kept as written (the cubic
			// ones use 3 u^2 and u_x u_y u_z without the factor rho)
			T = rT - rho * (vx * vx + vy * vy + vz * vz);
			Nxz = rNxz + rho * (vz * vz - vx * vx);
			Nyz = rNyz + rho * (vz * vz - vy * vy);
			Pxy = M110 - rho * vx * vy;
			Pxz = M101 - rho * vx * vz;
			Pyz = M011 - rho * vy * vz;
			Qc[0] = M210 - n1o3 * (no6 * vx * Pxy + vy * (no3 * vx * vx + no2 * Nxz - Nyz + T));
			Qc[1] = M201 - n1o3 * (no6 * vx * Pxz + vz * (no3 * vx * vx + no2 * Nxz - Nyz + T));
			Qc[2] = M120 - n1o3 * (no6 * vy * Pxy + vx * (no3 * vy * vy + no2 * Nyz - Nxz + T));
			Qc[3] = M021 - n1o3 * (no6 * vy * Pyz + vz * (no3 * vy * vy + no2 * Nyz - Nxz + T));
			Qc[4] = M102 - n1o3 * (no6 * vz * Pxz + vx * (no3 * vz * vz - Nyz - Nxz + T));
			Qc[5] = M012 - n1o3 * (no6 * vz * Pyz + vy * (no3 * vz * vz - Nyz - Nxz + T));
			Qc[6] = M111 - vx * Pyz - vy * Pxz - vz * Pxy - vx * vy * vz;
			eT = rho * no3 * n1o3;
			eNxz = eNyz = ePxy = ePxz = ePyz = 0;
			for (int k = 0; k < 7; k++)
				eQc[k] = 0;
		}
		dreal Ds[27], Dh[27];
		for (int q = 0; q < 27; q++) {
			dreal s = tensD(q, Nxz, Nyz, Pxy, Pxz, Pyz, CENTRAL) - tensD(q, eNxz, eNyz, ePxy, ePxz, ePyz, CENTRAL);
			if constexpr (WT)
				s = s + tensT(q, T, CENTRAL) - tensT(q, eT, CENTRAL);
			if constexpr (WQ)
				s = s + tensQ(q, Qc, CENTRAL) - tensQ(q, eQc, CENTRAL);
			Ds[q] = s;
			Dh[q] = f[q] - feq[q] - s;
		}
		const dreal beta = no1 / (no2 * KS.lbmViscosity / n1o3 + no1);
		dreal num = 0, den = 0;
		for (int i = 0; i < 27; i++) {
			const int q = LEX[i];
			num = i == 0 ? Ds[q] * Dh[q] * ifeq[q] : num + Ds[q] * Dh[q] * ifeq[q];
			den = i == 0 ? Dh[q] * Dh[q] * ifeq[q] : den + Dh[q] * Dh[q] * ifeq[q];
		}
		const dreal gamma = no1 / beta - (no2 - no1 / beta) * num / den;
		const dreal F[3] = {KS.fx, KS.fy, KS.fz};
		for (int q = 0; q < 27; q++) {
			const int c[3] = {CX[q], CY[q], CZ[q]};
			dreal t[3];
			for (int d = 0; d < 3; d++)
				t[d] = c[d] == 0 ? -v[d] * F[d] : (c[d] > 0 ? (-v[d] + no1) * F[d] : (-v[d] - no1) * F[d]);
			const dreal S = (no3 * (t[0] + t[1] + t[2])) / rho;
			KS.f[q] -= beta * (no2 * Ds[q] + gamma * Dh[q]) - (no1 - beta) * S * feq[q];
		}
	}

	// every absolute-form operator by id
	static void collide_any(KS27<dreal>& KS, int coll)
	{
		switch (coll) {
			case TNL_LBM_COLL_CUM: cumulant<false, false>(KS); break;
			case TNL_LBM_COLL_CUM_2017: cumulant<true, false>(KS); break;
			case TNL_LBM_COLL_CUM_2017_ANTIALIAS: cumulant<true, true>(KS); break;
			case TNL_LBM_COLL_MRT_LES: mrt(KS); break;
			case TNL_LBM_COLL_KBC_N1: kbc<false, false, false>(KS); break;
			case TNL_LBM_COLL_KBC_N2: kbc<false, true, false>(KS); break;
			case TNL_LBM_COLL_KBC_N3: kbc<false, false, true>(KS); break;
			case TNL_LBM_COLL_KBC_N4: kbc<false, true, true>(KS); break;
			case TNL_LBM_COLL_KBC_C1: kbc<true, false, false>(KS); break;
			case TNL_LBM_COLL_KBC_C2: kbc<true, true, false>(KS); break;
			case TNL_LBM_COLL_KBC_C3: kbc<true, false, true>(KS); break;
			default: kbc<true, true, true>(KS); break;
		}
	}
	// equilibrium each operator uses for initialisation and boundaries
	// (the LBM_EQ template default of each plugin)
	static int eq_of(int coll)
	{
		if (coll <= TNL_LBM_COLL_CUM_2017_ANTIALIAS)
			return 0;
		if (coll == TNL_LBM_COLL_MRT_LES || coll == TNL_LBM_COLL_KBC_N1 || coll == TNL_LBM_COLL_KBC_C1)
			return 1;
		return 2;
	}

};

// ===========================================================================
// DEVIATION FORM (float arithmetic on g = f - w): the algebra of
// common_well.h:15-51, eq_well.h:21-134 and col_cum_well.h:13-487 for the
// default cumulant.  Every chain value is kept as its deviation from the
// value it takes at rest (rho = 1, u = 0); the rest values are products of
// the 1-D weight moments (1, 0, 1/3).
// ===========================================================================
template <typename dreal_>
struct DevT
{
	using dreal = dreal_;
	static constexpr dreal one = 1, two = 2, three = 3, half = (dreal) 0.5;
	static constexpr dreal third = (dreal) (1.0 / 3.0), ninth = (dreal) (1.0 / 9.0), n1o27 = (dreal) (1.0 / 27.0);

	// common_well.h:15-51: rho = 1 + sum g (same grouping as common.h)
	static void densityVelocity(KS27<dreal>& KS)
	{
		const dreal* f = KS.f;
		KS.rho = ((((f[ppp] + f[mmm]) + (f[pmp] + f[mpm])) + ((f[ppm] + f[mmp]) + (f[mpp] + f[pmm])))
				  + (((f[zpp] + f[zmm]) + (f[zpm] + f[zmp])) + ((f[pzp] + f[mzm]) + (f[pzm] + f[mzp]))
					 + ((f[ppz] + f[mmz]) + (f[pmz] + f[mpz])))
				  + ((f[pzz] + f[mzz]) + (f[zpz] + f[zmz]) + (f[zzp] + f[zzm])))
			   + f[zzz];
		KS.rho += one;
		KS.vx = ((((f[ppp] - f[mmm]) + (f[pmp] - f[mpm])) + ((f[ppm] - f[mmp]) + (f[pmm] - f[mpp])))
				 + (((f[pzp] - f[mzm]) + (f[pzm] - f[mzp])) + ((f[ppz] - f[mmz]) + (f[pmz] - f[mpz]))) + (f[pzz] - f[mzz])
				 + KS.fx * half)
			  / KS.rho;
		KS.vy = ((((f[ppp] - f[mmm]) + (f[ppm] - f[mmp])) + ((f[mpp] - f[pmm]) + (f[mpm] - f[pmp])))
				 + (((f[ppz] - f[mmz]) + (f[mpz] - f[pmz])) + ((f[zpp] - f[zmm]) + (f[zpm] - f[zmp]))) + (f[zpz] - f[zmz])
				 + KS.fy * half)
			  / KS.rho;
		KS.vz = ((((f[ppp] - f[mmm]) + (f[mpp] - f[pmm])) + ((f[pmp] - f[mpm]) + (f[mmp] - f[ppm])))
				 + (((f[zpp] - f[zmm]) + (f[zmp] - f[zpm])) + ((f[pzp] - f[mzm]) + (f[mzp] - f[pzm]))) + (f[zzp] - f[zzm])
				 + KS.fz * half)
			  / KS.rho;
	}

	// g_q = rho * prod W_d(v_d) - prod W_d(0), product-form (EQ_INV_CUM) equilibrium
	// in deviation form; W(0) = 2/3 - v^2, W(+-1) = 1/6 + (v^2 +- v)/2.
	static dreal w1(int c)
	{
		return c == 0 ? (dreal) (2.0 / 3.0) : (dreal) (1.0 / 6.0);
	}
	static dreal dw1(int c, dreal v)
	{
		return c == 0 ? -v * v : half * (v * v + (dreal) c * v);
	}
	static dreal eq_invcum(int q, dreal rho, dreal vx, dreal vy, dreal vz)
	{
		const dreal wx = w1(CX[q]), wy = w1(CY[q]), wz = w1(CZ[q]);
		const dreal dx = dw1(CX[q], vx), dy = dw1(CY[q], vy), dz = dw1(CZ[q], vz);
		const dreal d2 = dx * (wy + dy) + wx * dy;
		const dreal w2 = wx * wy;
		const dreal d3 = d2 * (wz + dz) + w2 * dz;
		const dreal w3 = w2 * wz;
		return d3 + (rho - one) * (w3 + d3);
	}
	// 2nd-order polynomial equilibrium (eq_well.h:15-134): w (rho feq - 1)
	static dreal eq_poly(int q, dreal rho, dreal vx, dreal vy, dreal vz)
	{
		const int nz = (CX[q] != 0) + (CY[q] != 0) + (CZ[q] != 0);
		const dreal w = nz == 0 ? (dreal) (8.0 / 27.0) : nz == 1 ? (dreal) (2.0 / 27.0) : nz == 2 ? (dreal) (1.0 / 54.0) : (dreal) (1.0 / 216.0);
		const dreal cu = CX[q] * vx + CY[q] * vy + CZ[q] * vz;
		const dreal d = -(dreal) 1.5 * (vx * vx + vy * vy + vz * vz) + three * cu + (dreal) 4.5 * cu * cu;
		return w * (d + (rho - one) * (one + d));
	}
	static dreal eq(int which, int q, dreal rho, dreal vx, dreal vy, dreal vz)
	{
		return which == 0 ? eq_invcum(q, rho, vx, vy, vz) : eq_poly(q, rho, vx, vy, vz);
	}
	static void setEquilibrium(KS27<dreal>& KS, int which)
	{
		for (int q = 0; q < 27; q++)
			KS.f[q] = eq(which, q, KS.rho, KS.vx, KS.vy, KS.vz);
	}

	// forward pass on deviations; R0 = rest value of the (full) sum
	static void fwd(dreal dm, dreal dz, dreal dp, dreal R0, dreal v, dreal v2, dreal& k0, dreal& k1, dreal& k2)
	{
		k0 = (dp + dm) + dz;
		k1 = (dp - dm) - v * (R0 + k0);
		k2 = (dp + dm) - two * v * (dp - dm) + v2 * (R0 + k0);
	}
	static void bwd(dreal k0, dreal k1, dreal k2, dreal R0, dreal v, dreal v2, dreal& dm, dreal& dz, dreal& dp)
	{
		dz = k0 * (one - v2) - two * v * k1 - k2 - R0 * v2;
		dm = ((R0 + k0) * (v2 - v) + k1 * (two * v - one) + k2) * half;
		dp = ((R0 + k0) * (v2 + v) + k1 * (two * v + one) + k2) * half;
	}
	// rest values of a chain element after passes over axes with orders i (M_0=1, M_1=0, M_2=1/3)
	static dreal M(int i)
	{
		return i == 0 ? one : i == 1 ? (dreal) 0 : third;
	}

	// default D3Q27_CUM in deviation form
	static void cumulant(KS27<dreal>& KS)
	{
		dreal* g = KS.f;
		const dreal rho = KS.rho;
		const dreal drho = rho - one;
		const dreal rho_inv = one / rho;
		const dreal vx = KS.vx, vy = KS.vy, vz = KS.vz;
		const dreal vx2 = vx * vx, vy2 = vy * vy, vz2 = vz * vz;
		dreal G[3][3][3], T[3][3][3], U[3][3][3], k[3][3][3];
		for (int q = 0; q < 27; q++)
			G[CX[q] + 1][CY[q] + 1][CZ[q] + 1] = g[q];
		for (int a = 0; a < 3; a++)
			for (int b = 0; b < 3; b++)
				fwd(G[a][b][0], G[a][b][1], G[a][b][2], w1(a - 1) * w1(b - 1), vz, vz2, T[a][b][0], T[a][b][1], T[a][b][2]);
		for (int a = 0; a < 3; a++)
			for (int j = 0; j < 3; j++)
				fwd(T[a][0][j], T[a][1][j], T[a][2][j], w1(a - 1) * M(j), vy, vy2, U[a][0][j], U[a][1][j], U[a][2][j]);
		for (int i = 0; i < 3; i++)
			for (int j = 0; j < 3; j++)
				fwd(U[0][i][j], U[1][i][j], U[2][i][j], M(i) * M(j), vx, vx2, k[0][i][j], k[1][i][j], k[2][i][j]);
		// k[][][] are deviations of the central moments from their rest values M(i)M(j)M(k)
		const dreal omega1 = one / (three * KS.lbmViscosity + half);
		const dreal omega2 = one;
		const dreal C_110 = k[1][1][0], C_101 = k[1][0][1], C_011 = k[0][1][1];
		const dreal C_200 = k[2][0][0], C_020 = k[0][2][0], C_002 = k[0][0][2];
		const dreal Eq33 = (one - omega1) * (C_200 - C_020);
		const dreal Eq34 = (one - omega1) * (C_200 - C_002);
		const dreal Eq35 = drho * omega2 + (one - omega2) * (C_200 + C_020 + C_002);
		const dreal a = third * (Eq33 + Eq34 + Eq35);        // ks_200 - 1/3
		const dreal b = third * (-two * Eq33 + Eq34 + Eq35);  // ks_020 - 1/3
		const dreal c = third * (Eq33 - two * Eq34 + Eq35);   // ks_002 - 1/3
		const dreal s = (one - omega1) * C_110, p = (one - omega1) * C_101, qq = (one - omega1) * C_011;
		// 4th order (rest 0 for 211-type, 1/9 for 220-type)
		const dreal ks_211 = ((third + a) * qq + two * p * s) * rho_inv;
		const dreal ks_121 = ((third + b) * p + two * s * qq) * rho_inv;
		const dreal ks_112 = ((third + c) * s + two * qq * p) * rho_inv;
		const dreal ks_220 = ((b * a + two * s * s) + (b + a) * third - drho * ninth) * rho_inv;
		const dreal ks_022 = ((c * b + two * qq * qq) + (c + b) * third - drho * ninth) * rho_inv;
		const dreal ks_202 = ((a * c + two * p * p) + (a + c) * third - drho * ninth) * rho_inv;
		// 6th order (rest 1/27)
		const dreal X1 = (a + b + c) * ninth + (ks_022 + ks_202 + ks_220) * third + a * ks_022 + b * ks_202 + c * ks_220
					   + (dreal) 4.0 * (qq * ks_211 + p * ks_121 + s * ks_112);
		const dreal X2 = (dreal) (2.0 / 3.0) * ((a + b + c) * third + (a * b + b * c + c * a) + two * (p * p + qq * qq + s * s))
					   + (dreal) 16.0 * s * p * qq + (dreal) 4.0 * (p * p * b + qq * qq * a + s * s * c) + two * a * b * c;
		const dreal ks_222 = X1 * rho_inv - X2 * rho_inv * rho_inv - drho * (drho - one) * n1o27 * rho_inv * rho_inv;
		dreal Ks[3][3][3] = {};
		Ks[0][0][0] = drho;  // k_000 deviation
		Ks[1][0][0] = -k[1][0][0];
		Ks[0][1][0] = -k[0][1][0];
		Ks[0][0][1] = -k[0][0][1];
		Ks[2][0][0] = a; Ks[0][2][0] = b; Ks[0][0][2] = c;
		Ks[1][1][0] = s; Ks[1][0][1] = p; Ks[0][1][1] = qq;
		Ks[2][1][1] = ks_211; Ks[1][2][1] = ks_121; Ks[1][1][2] = ks_112;
		Ks[2][2][0] = ks_220; Ks[0][2][2] = ks_022; Ks[2][0][2] = ks_202;
		Ks[2][2][2] = ks_222;
		for (int j = 0; j < 3; j++)
			for (int kk = 0; kk < 3; kk++)
				bwd(Ks[0][j][kk], Ks[1][j][kk], Ks[2][j][kk], M(j) * M(kk), vx, vx2, U[0][j][kk], U[1][j][kk], U[2][j][kk]);
		for (int a2 = 0; a2 < 3; a2++)
			for (int kk = 0; kk < 3; kk++)
				bwd(U[a2][0][kk], U[a2][1][kk], U[a2][2][kk], w1(a2 - 1) * M(kk), vy, vy2, T[a2][0][kk], T[a2][1][kk], T[a2][2][kk]);
		for (int a2 = 0; a2 < 3; a2++)
			for (int b2 = 0; b2 < 3; b2++)
				bwd(T[a2][b2][0], T[a2][b2][1], T[a2][b2][2], w1(a2 - 1) * w1(b2 - 1), vz, vz2, G[a2][b2][0], G[a2][b2][1], G[a2][b2][2]);
		for (int q = 0; q < 27; q++)
			g[q] = G[CX[q] + 1][CY[q] + 1][CZ[q] + 1];
	}
};

using Dev = DevT<float>;

// ===========================================================================
// Storage access
// ===========================================================================
template <int STORE>
struct Store;
template <>
struct Store<TNL_LBM_STORE_F64>
{
	using real = double;
	using elem = double;
	static constexpr bool dev = false;
	static real load(const elem* p, int64_t i, int) { return p[i]; }
	static void store(elem* p, int64_t i, int, real v) { p[i] = v; }
	static double to_abs(real v, int q) { return v; }
	static real from_abs(double v, int q) { return v; }
};
template <>
struct Store<TNL_LBM_STORE_F32>
{
	using real = float;
	using elem = float;
	static constexpr bool dev = false;
	static real load(const elem* p, int64_t i, int) { return p[i]; }
	static void store(elem* p, int64_t i, int, real v) { p[i] = v; }
	static double to_abs(real v, int q) { return v; }
	static real from_abs(double v, int q) { return (float) v; }
};
template <>
struct Store<TNL_LBM_STORE_F32_DEV>
{
	using real = float;
	using elem = float;
	static constexpr bool dev = true;
	static real load(const elem* p, int64_t i, int) { return p[i]; }
	static void store(elem* p, int64_t i, int, real v) { p[i] = v; }
	static double to_abs(real v, int q) { return (double) v + w27(q); }
	static real from_abs(double v, int q) { return (float) (v - w27(q)); }
};
template <>
struct Store<TNL_LBM_STORE_F16_DEV>
{
	using real = float;
	using elem = uint16_t;
	static constexpr bool dev = true;
	static real load(const elem* p, int64_t i, int) { return f16_to_f32(p[i]); }
	static void store(elem* p, int64_t i, int, real v) { p[i] = f32_to_f16(v); }
	static double to_abs(real v, int q) { return (double) v + w27(q); }
	static real from_abs(double v, int q) { return (float) (v - w27(q)); }
};
template <>
struct Store<TNL_LBM_STORE_BF16_DEV>
{
	using real = float;
	using elem = uint16_t;
	static constexpr bool dev = true;
	static real load(const elem* p, int64_t i, int) { return bf16_to_f32(p[i]); }
	static void store(elem* p, int64_t i, int, real v) { p[i] = f32_to_bf16(v); }
	static double to_abs(real v, int q) { return (double) v + w27(q); }
	static real from_abs(double v, int q) { return (float) (v - w27(q)); }
};

template <>
struct Store<TNL_LBM_STORE_I16_DEV>
{
	using real = float;
	using elem = int16_t;
	static constexpr bool dev = true;
	// decode: raw * (w_q * R/32767); encode: snorm16(v * (1/w_q) * (1/R))
	static int cls(int q) { return (CX[q] != 0) + (CY[q] != 0) + (CZ[q] != 0); }
	static real load(const elem* p, int64_t i, int q)
	{
		const float Rs = (float) (g_store_range[cls(q)] / 32767.0);
		return (float) p[i] * ((float) w27(q) * Rs);
	}
	static void store(elem* p, int64_t i, int q, real v)
	{
		const float invR = (float) (1.0 / g_store_range[cls(q)]);
		p[i] = f32_to_snorm16(v * ((float) (1.0 / w27(q)) * invR));
	}
	static double to_abs(real v, int q) { return (double) v + w27(q); }
	static real from_abs(double v, int q) { return (float) (v - w27(q)); }
};

// Physics selector: absolute restatement or deviation form
template <typename S>
using Phys = std::conditional_t<S::dev, Dev, Abs<typename S::real>>;

template <typename real>
void store_macro(void* macro, int mreal, int64_t XYZ, int64_t s, const KS27<real>& KS)
{
	if (! macro)
		return;
	if (mreal == TNL_LBM_REAL_F64) {
		double* m = (double*) macro;
		m[0 * XYZ + s] = KS.rho;
		m[1 * XYZ + s] = KS.vx;
		m[2 * XYZ + s] = KS.vy;
		m[3 * XYZ + s] = KS.vz;
	}
	else {
		float* m = (float*) macro;
		m[0 * XYZ + s] = (float) KS.rho;
		m[1 * XYZ + s] = (float) KS.vx;
		m[2 * XYZ + s] = (float) KS.vy;
		m[3 * XYZ + s] = (float) KS.vz;
	}
}

// ===========================================================================
// The fused per-site kernel: LBMKernel<NSE> (kernels.h:59-112) with
// D3Q27_BC_All (bc.h:61-422) and the chosen streaming pattern.
// ===========================================================================
template <int STORE, int COLL, int PATTERN>
struct Site27
{
	using S = Store<STORE>;
	using real = typename S::real;
	using elem = typename S::elem;
	using P = Phys<S>;
	static constexpr bool AA = PATTERN == TNL_LBM_PATTERN_AA;

	static int eq_kind()
	{
		return Abs<double>::eq_of(COLL);
	}
	// the operator's equilibrium (COLL::setEquilibrium, common.h:61-91)
	static void set_eq(KS27<real>& KS)
	{
		if constexpr (S::dev)
			P::setEquilibrium(KS, 0);
		else
			P::setEquilibriumAny(KS, eq_kind());
	}
	static real eqv(int q, real rho, real vx, real vy, real vz)
	{
		if constexpr (S::dev)
			return P::eq(0, q, rho, vx, vy, vz);
		else
			return P::eq_any(eq_kind(), q, rho, vx, vy, vz);
	}
	static void collide(KS27<real>& KS)
	{
		if constexpr (S::dev)
			P::cumulant(KS);
		else
			P::collide_any(KS, COLL);
	}

	// kernelInitIndices (kernels.h:5-57)
	static void indices(const Geometry& G, bool periodic, const bool dist[3], int64_t x, int64_t y, int64_t z,
						int64_t& xp, int64_t& xm, int64_t& yp, int64_t& ym, int64_t& zp, int64_t& zm)
	{
		if (periodic) {
			xp = (! dist[0] && x == G.X - 1) ? 0 : (x + 1);
			xm = (! dist[0] && x == 0) ? (G.X - 1) : (x - 1);
			yp = (! dist[1] && y == G.Y - 1) ? 0 : (y + 1);
			ym = (! dist[1] && y == 0) ? (G.Y - 1) : (y - 1);
			zp = (! dist[2] && z == G.Z - 1) ? 0 : (z + 1);
			zm = (! dist[2] && z == 0) ? (G.Z - 1) : (z - 1);
		}
		else if (AA) {
			// ghost layers are assumed (kernels.h:30-37); an index that would leave the
			// stored array wraps instead of reading outside the allocation
			xp = x + 1 < G.X + G.ox ? x + 1 : 0; xm = x - 1 >= -G.ox ? x - 1 : G.X - 1;
			yp = y + 1 < G.Y + G.oy ? y + 1 : 0; ym = y - 1 >= -G.oy ? y - 1 : G.Y - 1;
			zp = z + 1 < G.Z + G.oz ? z + 1 : 0; zm = z - 1 >= -G.oz ? z - 1 : G.Z - 1;
		}
		else {
			// with overlaps (MPI build) the clamp extends into them (kernels.h:38-47)
			xp = std::min<int64_t>(x + 1, G.X - 1 + G.ox); xm = std::max<int64_t>(x - 1, -G.ox);
			yp = std::min<int64_t>(y + 1, G.Y - 1 + G.oy); ym = std::max<int64_t>(y - 1, -G.oy);
			zp = std::min<int64_t>(z + 1, G.Z - 1 + G.oz); zm = std::max<int64_t>(z - 1, -G.oz);
		}
	}
	static int64_t nb(int c, int64_t m, int64_t o, int64_t p)
	{
		return c < 0 ? m : (c > 0 ? p : o);
	}

	// STREAMING::streaming (streaming_AA.h:79-116, streaming_AB.h:20-60)
	static void stream_in(const Geometry& G, const tnl_lbm_d3q27_data* SD, KS27<real>& KS, int64_t xm, int64_t x,
						  int64_t xp, int64_t ym, int64_t y, int64_t yp, int64_t zm, int64_t z, int64_t zp)
	{
		const elem* df = (const elem*) SD->df[0];
		if (AA && SD->even_iter) {
			const int64_t s = G.s(x, y, z);
			for (int i = 0; i < 27; i++)
				KS.f[i] = S::load(df, i * G.XYZ + s, i);
		}
		else if (AA) {
			// read from the upstream neighbour, opposite slot: KS.f[i] = df(opp(i), r - c_i)
			for (int i = 0; i < 27; i++) {
				const int64_t s = G.s(nb(-CX[i], xm, x, xp), nb(-CY[i], ym, y, yp), nb(-CZ[i], zm, z, zp));
				KS.f[i] = S::load(df, opp(i) * G.XYZ + s, opp(i));
			}
		}
		else {
			for (int i = 0; i < 27; i++) {
				const int64_t s = G.s(nb(-CX[i], xm, x, xp), nb(-CY[i], ym, y, yp), nb(-CZ[i], zm, z, zp));
				KS.f[i] = S::load(df, i * G.XYZ + s, i);
			}
		}
	}
	// STREAMING::postCollisionStreaming (streaming_AA.h:14-76, streaming_AB.h:14-19)
	static void stream_out(const Geometry& G, const tnl_lbm_d3q27_data* SD, const KS27<real>& KS, int64_t xm, int64_t x,
						   int64_t xp, int64_t ym, int64_t y, int64_t yp, int64_t zm, int64_t z, int64_t zp)
	{
		if (AA && SD->even_iter) {
			elem* df = (elem*) SD->df[0];
			const int64_t s = G.s(x, y, z);
			for (int i = 0; i < 27; i++)
				S::store(df, opp(i) * G.XYZ + s, opp(i), KS.f[i]);
		}
		else if (AA) {
			elem* df = (elem*) SD->df[0];
			for (int i = 0; i < 27; i++) {
				const int64_t s = G.s(nb(CX[i], xm, x, xp), nb(CY[i], ym, y, yp), nb(CZ[i], zm, z, zp));
				S::store(df, i * G.XYZ + s, i, KS.f[i]);
			}
		}
		else {
			elem* df = (elem*) SD->df[1];
			const int64_t s = G.s(x, y, z);
			for (int i = 0; i < 27; i++)
				S::store(df, i * G.XYZ + s, i, KS.f[i]);
		}
	}

	static void inflow(const tnl_lbm_d3q27_data* SD, KS27<real>& KS, int64_t y, int64_t z)
	{
		// NSE_Data_ConstInflow::inflow / NSE_Data_InflowProfile::inflow (lbm_data.h:85-132)
		const int64_t i = SD->Y * z + y;
		KS.vx = (real) inflow_at(SD, 0, i);
		KS.vy = (real) inflow_at(SD, 1, i);
		KS.vz = (real) inflow_at(SD, 2, i);
	}

	static bool supported_geo(int16_t g)
	{
		if (g == GEO_OUTFLOW_RIGHT || g == GEO_OUTFLOW_RIGHT_INTERP || g == GEO_INFLOW_BB_LEFT)
			return ! AA;  // these read upstream or unstreamed slots: defined for the pull scheme only
		if (S::dev && g == GEO_INFLOW_LEFT)
			return false;
		return true;
	}

	static void site(const Geometry& G, const tnl_lbm_d3q27_data* SD, const bool dist[3], bool output_macro, int mreal,
					 int64_t x, int64_t y, int64_t z)
	{
		const int16_t gi_map = SD->map ? SD->map[G.s(x, y, z)] : (int16_t) GEO_PERIODIC;
		int64_t xp, xm, yp, ym, zp, zm;
		indices(G, gi_map == GEO_PERIODIC, dist, x, y, z, xp, xm, yp, ym, zp, zm);
		KS27<real> KS;
		// D3Q27_MACRO_Default::copyQuantities (macro.h:80-86)
		KS.lbmViscosity = (real) SD->lbm_viscosity;
		KS.fx = (real) SD->fx;
		KS.fy = (real) SD->fy;
		KS.fz = (real) SD->fz;
		add_site_force(SD, mreal, G.s(x, y, z), KS.fx, KS.fy, KS.fz);
		pre_collision(G, SD, KS, gi_map, xm, x, xp, ym, y, yp, zm, z, zp);
		if (do_collision(gi_map))
			collide(KS);
		if (gi_map != GEO_NOTHING)
			stream_out(G, SD, KS, xm, x, xp, ym, y, yp, zm, z, zp);
		if (output_macro)
			store_macro(SD->macro, mreal, G.XYZ, G.s(x, y, z), KS);
	}

	// D3Q27_BC_All::doCollision (bc.h:406-412)
	static bool do_collision(int16_t g)
	{
		return g == GEO_FLUID || g == GEO_PERIODIC || g == GEO_OUTFLOW_RIGHT || g == GEO_OUTFLOW_RIGHT_INTERP
			|| g == GEO_INFLOW_LEFT || g == GEO_INFLOW_BB_LEFT;
	}

	static real unstreamed(const Geometry& G, const tnl_lbm_d3q27_data* SD, int q, int64_t x, int64_t y, int64_t z)
	{
		return S::load((const elem*) SD->df[0], q * G.XYZ + G.s(x, y, z), q);
	}

	// D3Q27_BC_All::preCollision (bc.h:61-404)
	static void pre_collision(const Geometry& G, const tnl_lbm_d3q27_data* SD, KS27<real>& KS, int16_t gi, int64_t xm,
							  int64_t x, int64_t xp, int64_t ym, int64_t y, int64_t yp, int64_t zm, int64_t z, int64_t zp)
	{
		constexpr real one = 1;
		if (gi == GEO_NOTHING) {
			KS.rho = 1;
			KS.vx = 0;
			KS.vy = 0;
			KS.vz = 0;
			return;
		}
		if (gi == GEO_OUTFLOW_RIGHT)
			xp = x = xm;
		real* f = KS.f;
		if (gi == GEO_OUTFLOW_RIGHT_INTERP) {
			// streamingInterpRight (streaming_AB.h:210-244)
			constexpr real SpeedOfSound = (real) 0.5773502691896257;
			for (int i = 0; i < 27; i++) {
				const int64_t yy = nb(-CY[i], ym, y, yp), zz = nb(-CZ[i], zm, z, zp);
				if (CX[i] < 0)
					f[i] = SpeedOfSound * unstreamed(G, SD, i, xm, yy, zz) + (1 - SpeedOfSound) * unstreamed(G, SD, i, x, yy, zz);
				else
					f[i] = unstreamed(G, SD, i, CX[i] > 0 ? xm : x, yy, zz);
			}
		}
		else
			stream_in(G, SD, KS, xm, x, xp, ym, y, yp, zm, z, zp);

		const real rest_sum = S::dev ? one : (real) 0;  // sum of w over the 9+9 known directions
		switch (gi) {
			case GEO_INFLOW:
				inflow(SD, KS, y, z);
				KS.rho = 1;
				set_eq(KS);
				break;
			case GEO_INFLOW_LEFT:
				{
					// moment-based inflow (bc.h:91-142)
					inflow(SD, KS, y, z);
					KS.rho = (real) 1.0 / (1 - KS.vx)
						   * ((f[zzz] + (+((f[zpp] + f[zmm]) + (f[zpm] + f[zmp])) + ((f[zpz] + f[zmz]) + (f[zzp] + f[zzm]))))
							  + 2 * (f[mzz] + (+((f[mpp] + f[mmm]) + (f[mpm] + f[mmp])) + ((f[mpz] + f[mmz]) + (f[mzp] + f[mzm])))));
					const real n1o3 = (real) (1.0 / 3.0), n1o9 = (real) (1.0 / 9.0);
					const real m100 = KS.rho * KS.vx, m010 = KS.rho * KS.vy, m001 = KS.rho * KS.vz;
					const real m011 = KS.rho * (KS.vy * KS.vz);
					const real m020 = n1o3 * KS.rho + KS.rho * (KS.vy * KS.vy);
					const real m002 = n1o3 * KS.rho + KS.rho * (KS.vz * KS.vz);
					const real m021 = n1o3 * KS.rho * KS.vz + KS.rho * ((KS.vy * KS.vy) * KS.vz);
					const real m012 = n1o3 * KS.rho * KS.vy + KS.rho * (KS.vy * (KS.vz * KS.vz));
					const real m022 = n1o9 * KS.rho + n1o3 * KS.rho * (KS.vy * KS.vy + KS.vz * KS.vz) + KS.rho * (KS.vy * KS.vy) * (KS.vz * KS.vz);
					f[pzz] = m100 + (m022 - (m020 + m002)) + f[mzz]
						   + (+((f[zpp] + f[zmm]) + (f[zpm] + f[zmp])) + ((f[zzp] + f[zzm]) + (f[zpz] + f[zmz])))
						   + 2 * (+((f[mpp] + f[mmm]) + (f[mpm] + f[mmp])) + ((f[mpz] + f[mmz]) + (f[mzp] + f[mzm])));
					f[ppz] = (real) 0.5 * ((m020 - m022) + (-m012 + m010)) - (f[mpz] + f[zpz]);
					f[pmz] = (real) 0.5 * ((m020 - m022) + (m012 - m010)) - (f[mmz] + f[zmz]);
					f[pzp] = (real) 0.5 * ((m002 - m022) + (-m021 + m001)) - (f[mzp] + f[zzp]);
					f[pzm] = (real) 0.5 * ((m002 - m022) + (m021 - m001)) - (f[mzm] + f[zzm]);
					f[ppp] = (real) 0.25 * ((m022 + m011) + (m021 + m012)) - (f[mpp] + f[zpp]);
					f[ppm] = (real) 0.25 * ((m022 - m011) + (-m021 + m012)) - (f[mpm] + f[zpm]);
					f[pmp] = (real) 0.25 * ((m022 - m011) + (m021 - m012)) - (f[mmp] + f[zmp]);
					f[pmm] = (real) 0.25 * ((m022 + m011) + (-m021 - m012)) - (f[mmm] + f[zmm]);
					break;
				}
			case GEO_INFLOW_BB_LEFT:
				{
					// bc.h:143-165 (pull scheme only)
					inflow(SD, KS, y, z);
					const real no2 = 2, no3 = 3, n2o27 = (real) (2.0 / 27.0), n1o54 = (real) (1.0 / 54.0), n1o216 = (real) (1.0 / 216.0);
					f[pzz] = unstreamed(G, SD, mzz, x, y, z) - no2 * n2o27 * KS.rho * no3 * (-KS.vx);
					f[ppz] = unstreamed(G, SD, mmz, x, y, z) - no2 * n1o54 * KS.rho * no3 * (-KS.vx - KS.vy);
					f[pmz] = unstreamed(G, SD, mpz, x, y, z) - no2 * n1o54 * KS.rho * no3 * (-KS.vx + KS.vy);
					f[pzp] = unstreamed(G, SD, mzm, x, y, z) - no2 * n1o54 * KS.rho * no3 * (-KS.vx - KS.vz);
					f[pzm] = unstreamed(G, SD, mzp, x, y, z) - no2 * n1o54 * KS.rho * no3 * (-KS.vx + KS.vz);
					f[ppp] = unstreamed(G, SD, mmm, x, y, z) - no2 * n1o216 * KS.rho * no3 * (-KS.vx - KS.vy - KS.vz);
					f[ppm] = unstreamed(G, SD, mmp, x, y, z) - no2 * n1o216 * KS.rho * no3 * (-KS.vx - KS.vy + KS.vz);
					f[pmp] = unstreamed(G, SD, mpm, x, y, z) - no2 * n1o216 * KS.rho * no3 * (-KS.vx + KS.vy - KS.vz);
					f[pmm] = unstreamed(G, SD, mpp, x, y, z) - no2 * n1o216 * KS.rho * no3 * (-KS.vx + KS.vy + KS.vz);
					P::densityVelocity(KS);
					break;
				}
			case GEO_INFLOW_EQ_LEFT:
				// bc.h:166-185; with deviation storage the known-direction weights sum to 1
				inflow(SD, KS, y, z);
				KS.rho = (real) 1.0 / (1 - KS.vx)
					   * ((rest_sum + (f[zzz] + (+((f[zpp] + f[zmm]) + (f[zpm] + f[zmp])) + ((f[zpz] + f[zmz]) + (f[zzp] + f[zzm])))))
						  + 2 * (f[mzz] + (+((f[mpp] + f[mmm]) + (f[mpm] + f[mmp])) + ((f[mpz] + f[mmz]) + (f[mzp] + f[mzm])))));
				set_eq(KS);
				break;
			case GEO_OUTFLOW_EQ:
				P::densityVelocity(KS);
				KS.rho = 1;
				set_eq(KS);
				break;
			case GEO_OUTFLOW_RIGHT:
				P::densityVelocity(KS);
				KS.rho = 1;
				break;
			case GEO_OUTFLOW_RIGHT_INTERP:
				{
					// bc.h:195-200 with setEquilibriumDecomposition (common.h:94-124)
					P::densityVelocity(KS);
					for (int q = 0; q < 27; q++)
						f[q] += eqv(q, (real) 1, KS.vx, KS.vy, KS.vz) - eqv(q, KS.rho, KS.vx, KS.vy, KS.vz);
					KS.rho = 1;
					break;
				}
			case GEO_WALL:
				// bc.h:201-221: full-way bounce-back, no collision
				KS.rho = 1;
				KS.vx = 0;
				KS.vy = 0;
				KS.vz = 0;
				for (int q = 1; q < 27; q += 2) {
					const real t = f[q];
					f[q] = f[q + 1];
					f[q + 1] = t;
				}
				break;
			case GEO_SYM_TOP:
			case GEO_SYM_BOTTOM:
			case GEO_SYM_LEFT:
			case GEO_SYM_RIGHT:
			case GEO_SYM_BACK:
			case GEO_SYM_FRONT:
				{
					// bc.h:222-293: mirror the populations that point out of the symmetry plane
					int axis, sgn;
					switch (gi) {
						case GEO_SYM_TOP: axis = 2; sgn = -1; break;
						case GEO_SYM_BOTTOM: axis = 2; sgn = 1; break;
						case GEO_SYM_LEFT: axis = 0; sgn = 1; break;
						case GEO_SYM_RIGHT: axis = 0; sgn = -1; break;
						case GEO_SYM_BACK: axis = 1; sgn = 1; break;
						default: axis = 1; sgn = -1; break;
					}
					for (int q = 0; q < 27; q++) {
						const int c[3] = {CX[q], CY[q], CZ[q]};
						if (c[axis] == sgn) {
							int m[3] = {c[0], c[1], c[2]};
							m[axis] = -sgn;
							f[q] = f[dir_of(m[0], m[1], m[2])];
						}
					}
					P::densityVelocity(KS);
					break;
				}
			default:
				P::densityVelocity(KS);
				break;
		}
	}

	static int run(const tnl_lbm_d3q27_data* SD, const int64_t* off, const int64_t* end, uint32_t flags, int mreal, int nthreads)
	{
		const Geometry G{SD->X, SD->Y, SD->Z, SD->ox, SD->oy, SD->oz, SD->XYZ};
		const bool dist[3] = {(flags & TNL_LBM_FLAG_DISTRIBUTED_X) != 0, (flags & TNL_LBM_FLAG_DISTRIBUTED_Y) != 0,
							  (flags & TNL_LBM_FLAG_DISTRIBUTED_Z) != 0};
		const bool om = (flags & TNL_LBM_FLAG_OUTPUT_MACRO) != 0;
		if (SD->map) {
			for (int64_t x = off[0]; x < end[0]; x++)
				for (int64_t z = off[2]; z < end[2]; z++)
					for (int64_t y = off[1]; y < end[1]; y++)
						if (! supported_geo(SD->map[G.s(x, y, z)]))
							return fail(TNL_LBM_ERR_UNSUPPORTED, "geometry code not supported with this pattern/storage");
		}
		// state.hpp:1388-1397: collapse(2) over x,z with y innermost
#pragma omp parallel for schedule(static) collapse(2) num_threads(nthreads > 0 ? nthreads : 1)
		for (int64_t x = off[0]; x < end[0]; x++)
			for (int64_t z = off[2]; z < end[2]; z++)
				for (int64_t y = off[1]; y < end[1]; y++)
					site(G, SD, dist, om, mreal, x, y, z);
		return TNL_LBM_OK;
	}
};

// ===========================================================================
// D3Q7 advection-diffusion (d3q7/*.h)
// ===========================================================================
constexpr int Q7_OPP[7] = {0, 2, 1, 4, 3, 6, 5};
double w7(int q)
{
	return q == 0 ? 0.25 : 0.125;
}

template <typename real>
struct Ade
{
	static constexpr real iCs2 = 4, w_0 = (real) 0.25, w_1 = (real) 0.125;
	static constexpr real no1 = 1, no2 = 2, n1o2 = (real) 0.5;
	// D3Q7_EQ::feq / eq_* (d3q7/eq.h:19-53)
	static real feq(int qx, int qy, int qz, real vx, real vy, real vz)
	{
		return no1 - n1o2 * iCs2 * (vx * vx + vy * vy + vz * vz) + iCs2 * (qx * vx + qy * vy + qz * vz)
			 + n1o2 * iCs2 * iCs2 * (qx * vx + qy * vy + qz * vz) * (qx * vx + qy * vy + qz * vz);
	}
	static real eq(int q, real phi, real vx, real vy, real vz)
	{
		return (q == 0 ? w_0 : w_1) * phi * feq(CX[q], CY[q], CZ[q], vx, vy, vz);
	}
	static void setEquilibrium(KS7<real>& KS)
	{
		for (int q = 0; q < 7; q++)
			KS.f[q] = eq(q, KS.phi, KS.vx, KS.vy, KS.vz);
	}
	// D3Q7_COMMON::computeDensityAndVelocity (d3q7/common.h:16-25)
	static void density(KS7<real>& KS)
	{
		real phi = 0;
		for (int i = 0; i < 7; i++)
			phi += KS.f[i];
		KS.phi = phi;
	}
	// D3Q7_SRT (d3q7/col_srt.h:13-26)
	static void srt(KS7<real>& KS)
	{
		const real tau = n1o2 + iCs2 * KS.lbmViscosity;
		const real omega = no1 / tau;
		const int order[7] = {zzz, pzz, zpz, zzp, mzz, zmz, zzm};
		for (int k = 0; k < 7; k++) {
			const int q = order[k];
			KS.f[q] += omega * (eq(q, KS.phi, KS.vx, KS.vy, KS.vz) - KS.f[q]);
		}
	}
	// D3Q7_MRT (d3q7/col_mrt.h:15-50)
	static void mrt(KS7<real>& KS)
	{
		const real Cs2 = no1 / iCs2;
		const real omega = no1 / (n1o2 + iCs2 * KS.lbmViscosity);
		const real omega2 = omega, omega3 = omega, omega4 = omega, omega5 = no1, omega6 = no1, omega7 = no1;
		real* f = KS.f;
		real m_100 = KS.phi * KS.vx + f[mzz] - f[pzz];
		real m_010 = KS.phi * KS.vy + f[zmz] - f[zpz];
		real m_001 = KS.phi * KS.vz + f[zzm] - f[zzp];
		real m_200 = KS.phi * (KS.vx * KS.vx + Cs2) - f[mzz] - f[pzz];
		real m_020 = KS.phi * (KS.vy * KS.vy + Cs2) - f[zmz] - f[zpz];
		real m_002 = KS.phi * (KS.vz * KS.vz + Cs2) - f[zzm] - f[zzp];
		m_100 *= omega2; m_010 *= omega3; m_001 *= omega4; m_200 *= omega5; m_020 *= omega6; m_002 *= omega7;
		f[zzz] += -m_200 - m_020 - m_002;
		f[pzz] += n1o2 * (m_200 + m_100);
		f[zpz] += n1o2 * (m_020 + m_010);
		f[zzp] += n1o2 * (m_002 + m_001);
		f[mzz] += n1o2 * (m_200 - m_100);
		f[zmz] += n1o2 * (m_020 - m_010);
		f[zzm] += n1o2 * (m_002 - m_001);
	}
	// D3Q7_CLBM (d3q7/col_clbm.h:15-84)
	static void clbm(KS7<real>& KS)
	{
		const real Cs2 = no1 / iCs2;
		const real omega = no1 / (n1o2 + iCs2 * KS.lbmViscosity);
		const real omega2 = omega, omega3 = omega, omega4 = omega, omega5 = no1, omega6 = omega5, omega7 = omega5;
		real* f = KS.f;
		real k100 = KS.phi * KS.vx + f[mzz] - f[pzz];
		real k010 = KS.phi * KS.vy + f[zmz] - f[zpz];
		real k001 = KS.phi * KS.vz + f[zzm] - f[zzp];
		real k200 = KS.phi * (Cs2 - KS.vx * KS.vx) + no2 * KS.vx * (f[pzz] - f[mzz]) - f[mzz] - f[pzz];
		real k020 = KS.phi * (Cs2 - KS.vy * KS.vy) + no2 * KS.vy * (f[zpz] - f[zmz]) - f[zmz] - f[zpz];
		real k002 = KS.phi * (Cs2 - KS.vz * KS.vz) + no2 * KS.vz * (f[zzp] - f[zzm]) - f[zzm] - f[zzp];
		k100 *= omega2; k010 *= omega3; k001 *= omega4; k200 *= omega5; k020 *= omega6; k002 *= omega7;
		f[zzz] += -no2 * (k100 * KS.vx + k010 * KS.vy + k001 * KS.vz) - k200 - k020 - k002;
		f[pzz] += k100 * KS.vx + n1o2 * (k200 + k100);
		f[zpz] += k010 * KS.vy + n1o2 * (k020 + k010);
		f[zzp] += k001 * KS.vz + n1o2 * (k002 + k001);
		f[mzz] += k100 * KS.vx + n1o2 * (k200 - k100);
		f[zmz] += k010 * KS.vy + n1o2 * (k020 - k010);
		f[zzm] += k001 * KS.vz + n1o2 * (k002 - k001);
	}
};

// D3Q7 storage: absolute for F64/F32, (f - w_q) deviation otherwise,
// using the same element types as the D3Q27 array
template <int STORE>
struct Store7
{
	using S = Store<STORE>;
	using real = typename S::real;
	using elem = typename S::elem;
	// deviation forms: stored = (f_q - w_q phi_ref) * scale
	static real load(const elem* p, int64_t i, int q, real phi_ref, real scale)
	{
		if constexpr (STORE == TNL_LBM_STORE_I16_DEV)
			return ((float) p[i] * (float) (w7(q) / 32767.0)) / scale + (real) w7(q) * phi_ref;
		const real v = S::load(p, i, q);
		return S::dev ? v / scale + (real) w7(q) * phi_ref : v;
	}
	static void store(elem* p, int64_t i, int q, real v, real phi_ref, real scale)
	{
		if constexpr (STORE == TNL_LBM_STORE_I16_DEV)
			p[i] = f32_to_snorm16((v - (real) w7(q) * phi_ref) * scale * (float) (1.0 / w7(q)));
		else
			S::store(p, i, q, S::dev ? (v - (real) w7(q) * phi_ref) * scale : v);
	}
};

template <int STORE, int PATTERN>
struct Site7
{
	using S7 = Store7<STORE>;
	using real = typename S7::real;
	using elem = typename S7::elem;
	using A = Ade<real>;
	static constexpr bool AA = PATTERN == TNL_LBM_PATTERN_AA;

	static int64_t nb(int c, int64_t m, int64_t o, int64_t p)
	{
		return c < 0 ? m : (c > 0 ? p : o);
	}
	static void stream_in(const Geometry& G, const tnl_lbm_d3q7_data* SD, KS7<real>& KS, int64_t xm, int64_t x,
						  int64_t xp, int64_t ym, int64_t y, int64_t yp, int64_t zm, int64_t z, int64_t zp)
	{
		const elem* df = (const elem*) SD->df[0];
		const real pr = (real) SD->phi_ref, sc = (real) SD->store_scale;
		for (int i = 0; i < 7; i++) {
			if (AA && SD->even_iter)
				KS.f[i] = S7::load(df, i * G.XYZ + G.s(x, y, z), i, pr, sc);
			else if (AA)
				KS.f[i] = S7::load(df, Q7_OPP[i] * G.XYZ + G.s(nb(-CX[i], xm, x, xp), nb(-CY[i], ym, y, yp), nb(-CZ[i], zm, z, zp)), Q7_OPP[i], pr, sc);
			else
				KS.f[i] = S7::load(df, i * G.XYZ + G.s(nb(-CX[i], xm, x, xp), nb(-CY[i], ym, y, yp), nb(-CZ[i], zm, z, zp)), i, pr, sc);
		}
	}
	static void stream_out(const Geometry& G, const tnl_lbm_d3q7_data* SD, const KS7<real>& KS, int64_t xm, int64_t x,
						   int64_t xp, int64_t ym, int64_t y, int64_t yp, int64_t zm, int64_t z, int64_t zp)
	{
		const real pr = (real) SD->phi_ref, sc = (real) SD->store_scale;
		for (int i = 0; i < 7; i++) {
			if (AA && SD->even_iter)
				S7::store((elem*) SD->df[0], Q7_OPP[i] * G.XYZ + G.s(x, y, z), Q7_OPP[i], KS.f[i], pr, sc);
			else if (AA)
				S7::store((elem*) SD->df[0], i * G.XYZ + G.s(nb(CX[i], xm, x, xp), nb(CY[i], ym, y, yp), nb(CZ[i], zm, z, zp)), i, KS.f[i], pr, sc);
			else
				S7::store((elem*) SD->df[1], i * G.XYZ + G.s(x, y, z), i, KS.f[i], pr, sc);
		}
	}
	static real macro_at(const void* p, int mreal, int64_t i)
	{
		return mreal == TNL_LBM_REAL_F64 ? (real) ((const double*) p)[i] : (real) ((const float*) p)[i];
	}

	// ADE half of cudaLBMKernel<NSE,ADE> (kernels.h:172-195) with D3Q7_BC_All (d3q7/bc.h:61-212)
	static void site(const Geometry& G, const tnl_lbm_d3q7_data* SD, int mreal, bool output_macro, real vx, real vy,
					 real vz, int64_t xm, int64_t x, int64_t xp, int64_t ym, int64_t y, int64_t yp, int64_t zm, int64_t z,
					 int64_t zp)
	{
		const int64_t s = G.s(x, y, z);
		const int16_t gi = SD->map ? SD->map[s] : (int16_t) A_PERIODIC;
		KS7<real> KS;
		KS.vx = vx;
		KS.vy = vy;
		KS.vz = vz;
		// D3Q7_MACRO_Default::copyQuantities / ADE_Data::diffusionCoefficient (lbm_data.h:231-239)
		KS.lbmViscosity = SD->diffusion_coefficient ? macro_at(SD->diffusion_coefficient, mreal, s) : (real) SD->lbm_diffusivity;
		if (gi == A_NOTHING) {
			KS.phi = 0;
		}
		else {
			int64_t xx = x, xxp = xp;
			if (gi == A_OUTFLOW_RIGHT)
				xxp = xx = xm;
			stream_in(G, SD, KS, xm, xx, xxp, ym, y, yp, zm, z, zp);
			switch (gi) {
				case A_INFLOW:
					KS.phi = (real) SD->inflow_phi;
					A::setEquilibrium(KS);
					break;
				case A_OUTFLOW_RIGHT:
					stream_in(G, SD, KS, xm, xx, xm, ym, y, yp, zm, z, zp);
					A::density(KS);
					break;
				case A_WALL:
					A::density(KS);
					for (int q = 1; q < 7; q += 2) {
						const real t = KS.f[q];
						KS.f[q] = KS.f[q + 1];
						KS.f[q + 1] = t;
					}
					break;
				case A_SYM_TOP: KS.f[zzm] = KS.f[zzp]; A::density(KS); break;
				case A_SYM_BOTTOM: KS.f[zzp] = KS.f[zzm]; A::density(KS); break;
				case A_SYM_LEFT: KS.f[pzz] = KS.f[mzz]; A::density(KS); break;
				case A_SYM_RIGHT: KS.f[mzz] = KS.f[pzz]; A::density(KS); break;
				case A_SYM_BACK: KS.f[zpz] = KS.f[zmz]; A::density(KS); break;
				case A_SYM_FRONT: KS.f[zmz] = KS.f[zpz]; A::density(KS); break;
				default: A::density(KS); break;
			}
			const bool collide = gi == A_FLUID || gi == A_PERIODIC || gi == A_SOLID || gi == A_OUTFLOW_RIGHT;
			if (collide) {
				if (SD->collision == TNL_LBM_ADE_COLL_SRT)
					A::srt(KS);
				else if (SD->collision == TNL_LBM_ADE_COLL_MRT)
					A::mrt(KS);
				else
					A::clbm(KS);
				// source term (not in the reference, lbm_data.h:229 TODO): f_q += w_q S
				if (SD->source) {
					const real S = macro_at(SD->source, mreal, s);
					for (int q = 0; q < 7; q++)
						KS.f[q] += (real) w7(q) * S;
				}
			}
			stream_out(G, SD, KS, xm, x, xp, ym, y, yp, zm, z, zp);
		}
		if (output_macro && SD->macro) {
			if (mreal == TNL_LBM_REAL_F64)
				((double*) SD->macro)[s] = KS.phi;
			else
				((float*) SD->macro)[s] = (float) KS.phi;
		}
	}
};

template <int STORE, int COLL, int PATTERN>
int run_coupled(const tnl_lbm_d3q27_data* SD, const tnl_lbm_d3q7_data* ADE, const int64_t* off, const int64_t* end,
				uint32_t flags, int mreal, int nthreads)
{
	using N = Site27<STORE, COLL, PATTERN>;
	using A = Site7<STORE, PATTERN>;
	const Geometry G{SD->X, SD->Y, SD->Z, SD->ox, SD->oy, SD->oz, SD->XYZ};
	const Geometry GA{ADE->X, ADE->Y, ADE->Z, ADE->ox, ADE->oy, ADE->oz, ADE->XYZ};
	const bool dist[3] = {(flags & TNL_LBM_FLAG_DISTRIBUTED_X) != 0, (flags & TNL_LBM_FLAG_DISTRIBUTED_Y) != 0,
						  (flags & TNL_LBM_FLAG_DISTRIBUTED_Z) != 0};
	const bool om = (flags & TNL_LBM_FLAG_OUTPUT_MACRO) != 0;
	// state_NSE_ADE.h:304-309 runs this loop serially; the sites are independent, so it is parallel here
#pragma omp parallel for schedule(static) collapse(2) num_threads(nthreads > 0 ? nthreads : 1)
	for (int64_t x = off[0]; x < end[0]; x++)
		for (int64_t z = off[2]; z < end[2]; z++)
			for (int64_t y = off[1]; y < end[1]; y++) {
				const int64_t s = G.s(x, y, z);
				const int16_t gi = SD->map ? SD->map[s] : (int16_t) GEO_PERIODIC;
				int64_t xp, xm, yp, ym, zp, zm;
				N::indices(G, gi == GEO_PERIODIC, dist, x, y, z, xp, xm, yp, ym, zp, zm);
				KS27<typename N::real> KS;
				KS.lbmViscosity = (typename N::real) SD->lbm_viscosity;
				KS.fx = (typename N::real) SD->fx;
				KS.fy = (typename N::real) SD->fy;
				KS.fz = (typename N::real) SD->fz;
				add_site_force(SD, mreal, s, KS.fx, KS.fy, KS.fz);
				N::pre_collision(G, SD, KS, gi, xm, x, xp, ym, y, yp, zm, z, zp);
				if (N::do_collision(gi))
					N::collide(KS);
				if (gi != GEO_NOTHING)
					N::stream_out(G, SD, KS, xm, x, xp, ym, y, yp, zm, z, zp);
				if (om)
					store_macro(SD->macro, mreal, G.XYZ, s, KS);
				A::site(GA, ADE, mreal, om, KS.vx, KS.vy, KS.vz, xm, x, xp, ym, y, yp, zm, z, zp);
			}
	return TNL_LBM_OK;
}

// ---------------------------------------------------------------------------
int check(const tnl_lbm_config* cfg, const tnl_lbm_d3q27_data* SD)
{
	if (! cfg || ! SD)
		return fail(TNL_LBM_ERR_NULL, "null config or data");
	if (cfg->struct_size != sizeof(tnl_lbm_config) || SD->struct_size != sizeof(tnl_lbm_d3q27_data)
		|| cfg->version != TNL_LBM_ABI_VERSION || SD->version != TNL_LBM_ABI_VERSION)
		return fail(TNL_LBM_ERR_VERSION, "struct_size/version mismatch");
	if (SD->X <= 0 || SD->Y <= 0 || SD->Z <= 0 || SD->ox < 0 || SD->oy < 0 || SD->oz < 0)
		return fail(TNL_LBM_ERR_DIMS, "bad sizes");
	if (SD->XYZ < (SD->X + 2 * SD->ox) * (SD->Y + 2 * SD->oy) * (SD->Z + 2 * SD->oz))
		return fail(TNL_LBM_ERR_DIMS, "XYZ smaller than the stored lattice");
	if (! SD->df[0] || (cfg->pattern == TNL_LBM_PATTERN_AB && ! SD->df[1]))
		return fail(TNL_LBM_ERR_NULL, "population array missing");
	if (cfg->storage < 0 || cfg->storage > 5 || cfg->pattern < 0 || cfg->pattern > 1)
		return fail(TNL_LBM_ERR_ARG, "bad storage or pattern");
	for (int n = 0; n < 4; n++) {
		const double R = SD->store_range_class[n] > 0 ? SD->store_range_class[n] : SD->store_range;
		if (cfg->storage == TNL_LBM_STORE_I16_DEV && ! (R > 0))
			return fail(TNL_LBM_ERR_ARG, "store_range must be positive");
		g_store_range[n] = R;
	}
	if (cfg->collision < TNL_LBM_COLL_CUM || cfg->collision > TNL_LBM_COLL_KBC_C4)
		return fail(TNL_LBM_ERR_UNSUPPORTED, "unknown collision operator");
	const bool dev = cfg->storage >= TNL_LBM_STORE_F32_DEV;
	if (dev && cfg->collision != TNL_LBM_COLL_CUM)
		return fail(TNL_LBM_ERR_UNSUPPORTED, "deviation storage supports the default cumulant only");
	return TNL_LBM_OK;
}

template <int STORE, int PATTERN>
int dispatch_coll(const tnl_lbm_config* cfg, const tnl_lbm_d3q27_data* SD, const int64_t* off, const int64_t* end,
				  uint32_t flags, int nthreads)
{
#define RUN(C) return Site27<STORE, C, PATTERN>::run(SD, off, end, flags, cfg->macro_real, nthreads)
	if constexpr (Store<STORE>::dev) {
		RUN(TNL_LBM_COLL_CUM);  // check() admits only the default cumulant here
	}
	else {
		switch (cfg->collision) {
			case TNL_LBM_COLL_CUM: RUN(TNL_LBM_COLL_CUM);
			case TNL_LBM_COLL_CUM_2017: RUN(TNL_LBM_COLL_CUM_2017);
			case TNL_LBM_COLL_CUM_2017_ANTIALIAS: RUN(TNL_LBM_COLL_CUM_2017_ANTIALIAS);
			case TNL_LBM_COLL_MRT_LES: RUN(TNL_LBM_COLL_MRT_LES);
			case TNL_LBM_COLL_KBC_N1: RUN(TNL_LBM_COLL_KBC_N1);
			case TNL_LBM_COLL_KBC_N2: RUN(TNL_LBM_COLL_KBC_N2);
			case TNL_LBM_COLL_KBC_N3: RUN(TNL_LBM_COLL_KBC_N3);
			case TNL_LBM_COLL_KBC_N4: RUN(TNL_LBM_COLL_KBC_N4);
			case TNL_LBM_COLL_KBC_C1: RUN(TNL_LBM_COLL_KBC_C1);
			case TNL_LBM_COLL_KBC_C2: RUN(TNL_LBM_COLL_KBC_C2);
			case TNL_LBM_COLL_KBC_C3: RUN(TNL_LBM_COLL_KBC_C3);
			default: RUN(TNL_LBM_COLL_KBC_C4);
		}
	}
#undef RUN
}
template <int STORE>
int dispatch_pattern(const tnl_lbm_config* cfg, const tnl_lbm_d3q27_data* SD, const int64_t* off, const int64_t* end,
					 uint32_t flags, int nthreads)
{
	if (cfg->pattern == TNL_LBM_PATTERN_AA)
		return dispatch_coll<STORE, TNL_LBM_PATTERN_AA>(cfg, SD, off, end, flags, nthreads);
	return dispatch_coll<STORE, TNL_LBM_PATTERN_AB>(cfg, SD, off, end, flags, nthreads);
}

template <int STORE>
void set_eq_site(const tnl_lbm_config* cfg, const tnl_lbm_d3q27_data* SD, int64_t s, double rho, double vx, double vy, double vz)
{
	using S = Store<STORE>;
	using P = Phys<S>;
	using real = typename S::real;
	const int which = Abs<double>::eq_of(cfg->collision);
	for (int q = 0; q < 27; q++) {
		real v;
		if constexpr (S::dev)
			v = P::eq(0, q, (real) rho, (real) vx, (real) vy, (real) vz);
		else
			v = P::eq_any(which, q, (real) rho, (real) vx, (real) vy, (real) vz);
		for (int k = 0; k < (cfg->pattern == TNL_LBM_PATTERN_AB ? 2 : 1); k++)
			S::store((typename S::elem*) SD->df[k], q * SD->XYZ + s, q, v);
	}
}

template <int STORE>
int set_eq_all(const tnl_lbm_config* cfg, const tnl_lbm_d3q27_data* SD, const double* rho, const double* vx,
			   const double* vy, const double* vz, double r0, double x0, double y0, double z0)
{
	const int64_t NS = (SD->X + 2 * SD->ox) * (SD->Y + 2 * SD->oy) * (SD->Z + 2 * SD->oz);
	for (int64_t s = 0; s < NS; s++) {
		if (rho)
			set_eq_site<STORE>(cfg, SD, s, rho[s], vx[s], vy[s], vz[s]);
		else
			set_eq_site<STORE>(cfg, SD, s, r0, x0, y0, z0);
	}
	return TNL_LBM_OK;
}

template <int STORE>
int compute_macro_all(const tnl_lbm_config* cfg, const tnl_lbm_d3q27_data* SD)
{
	using S = Store<STORE>;
	using P = Phys<S>;
	const Geometry G{SD->X, SD->Y, SD->Z, SD->ox, SD->oy, SD->oz, SD->XYZ};
	for (int64_t x = 0; x < SD->X; x++)
		for (int64_t z = 0; z < SD->Z; z++)
			for (int64_t y = 0; y < SD->Y; y++) {
				// lbm_block.hpp:293-303 (forces zeroed by zeroForcesInKS)
				KS27<typename S::real> KS;
				const int64_t s = G.s(x, y, z);
				for (int q = 0; q < 27; q++)
					KS.f[q] = S::load((const typename S::elem*) SD->df[0], q * SD->XYZ + s, q);
				KS.lbmViscosity = (typename S::real) SD->lbm_viscosity;
				P::densityVelocity(KS);
				store_macro(SD->macro, cfg->macro_real, SD->XYZ, s, KS);
			}
	return TNL_LBM_OK;
}

template <int STORE>
int export_all(const tnl_lbm_d3q27_data* SD, int which, double* out, int Q)
{
	using S = Store<STORE>;
	const int64_t NS = (SD->X + 2 * SD->ox) * (SD->Y + 2 * SD->oy) * (SD->Z + 2 * SD->oz);
	for (int q = 0; q < Q; q++)
		for (int64_t s = 0; s < NS; s++) {
			const double v = S::load((const typename S::elem*) SD->df[which], q * SD->XYZ + s, q);
			out[q * SD->XYZ + s] = S::dev ? v + (Q == 27 ? w27(q) : w7(q)) : v;
		}
	return TNL_LBM_OK;
}
template <int STORE>
int import_all(const tnl_lbm_d3q27_data* SD, int which, const double* in)
{
	using S = Store<STORE>;
	const int64_t NS = (SD->X + 2 * SD->ox) * (SD->Y + 2 * SD->oy) * (SD->Z + 2 * SD->oz);
	for (int q = 0; q < 27; q++)
		for (int64_t s = 0; s < NS; s++)
			S::store((typename S::elem*) SD->df[which], q * SD->XYZ + s, q, S::from_abs(in[q * SD->XYZ + s], q));
	return TNL_LBM_OK;
}

#define STORE_SWITCH(storage, CALL)                                     \
	switch (storage) {                                                  \
		case TNL_LBM_STORE_F64: { constexpr int ST = TNL_LBM_STORE_F64; return CALL; } \
		case TNL_LBM_STORE_F32: { constexpr int ST = TNL_LBM_STORE_F32; return CALL; } \
		case TNL_LBM_STORE_F32_DEV: { constexpr int ST = TNL_LBM_STORE_F32_DEV; return CALL; } \
		case TNL_LBM_STORE_F16_DEV: { constexpr int ST = TNL_LBM_STORE_F16_DEV; return CALL; } \
		case TNL_LBM_STORE_BF16_DEV: { constexpr int ST = TNL_LBM_STORE_BF16_DEV; return CALL; } \
		default: { constexpr int ST = TNL_LBM_STORE_I16_DEV; return CALL; } \
	}

}  // namespace

// ===========================================================================
// ===========================================================================
// ADJOINT of one A-B step with the default cumulant (the transposed Jacobian
// of the step map), computed by forward-mode automatic differentiation of
// the restated site map above: a reference independent of the product's
// hand-written reverse pass.  Supported geometry codes: FLUID, PERIODIC,
// WALL, INFLOW, OUTFLOW_EQ, OUTFLOW_RIGHT, NOTHING.
// ===========================================================================
namespace adj {

constexpr int ND = 30;  // tangents: 27 pulled populations + 3 inflow velocity components
struct Dual
{
	double v;
	double d[ND];
	constexpr Dual() : v(0), d{} {}
	constexpr Dual(double x) : v(x), d{} {}
	Dual& operator+=(const Dual& b);
	Dual& operator-=(const Dual& b);
	Dual& operator*=(const Dual& b);
};
inline Dual operator+(const Dual& a, const Dual& b)
{
	Dual r(a.v + b.v);
	for (int i = 0; i < ND; i++)
		r.d[i] = a.d[i] + b.d[i];
	return r;
}
inline Dual operator-(const Dual& a, const Dual& b)
{
	Dual r(a.v - b.v);
	for (int i = 0; i < ND; i++)
		r.d[i] = a.d[i] - b.d[i];
	return r;
}
inline Dual operator-(const Dual& a)
{
	Dual r(-a.v);
	for (int i = 0; i < ND; i++)
		r.d[i] = -a.d[i];
	return r;
}
inline Dual operator*(const Dual& a, const Dual& b)
{
	Dual r(a.v * b.v);
	for (int i = 0; i < ND; i++)
		r.d[i] = a.d[i] * b.v + a.v * b.d[i];
	return r;
}
inline Dual operator/(const Dual& a, const Dual& b)
{
	Dual r(a.v / b.v);
	for (int i = 0; i < ND; i++)
		r.d[i] = (a.d[i] * b.v - a.v * b.d[i]) / (b.v * b.v);
	return r;
}
inline Dual& Dual::operator+=(const Dual& b)
{
	return *this = *this + b;
}
inline Dual& Dual::operator-=(const Dual& b)
{
	return *this = *this - b;
}
inline Dual& Dual::operator*=(const Dual& b)
{
	return *this = *this * b;
}

// the A-B site map (pulled populations, inflow velocity) -> post-collision
// populations for the supported codes, as pre_collision / collide do it
template <typename T>
void site_map(int16_t gi, KS27<T>& KS, const T uin[3], double nu, double fx, double fy, double fz)
{
	using A = Abs<T>;
	KS.lbmViscosity = nu;
	KS.fx = fx;
	KS.fy = fy;
	KS.fz = fz;
	switch (gi) {
		case GEO_INFLOW:
			KS.vx = uin[0];
			KS.vy = uin[1];
			KS.vz = uin[2];
			KS.rho = 1;
			A::setEquilibrium(KS, 0);
			return;
		case GEO_OUTFLOW_EQ:
			A::densityVelocity(KS);
			KS.rho = 1;
			A::setEquilibrium(KS, 0);
			return;
		case GEO_WALL:
			for (int q = 1; q < 27; q += 2) {
				const T t = KS.f[q];
				KS.f[q] = KS.f[q + 1];
				KS.f[q + 1] = t;
			}
			return;
		case GEO_OUTFLOW_RIGHT:
			A::densityVelocity(KS);
			KS.rho = 1;
			A::template cumulant<false, false>(KS);
			return;
		default:  // FLUID, PERIODIC
			A::densityVelocity(KS);
			A::template cumulant<false, false>(KS);
			return;
	}
}

bool supported(int16_t g)
{
	return g == GEO_FLUID || g == GEO_PERIODIC || g == GEO_WALL || g == GEO_INFLOW || g == GEO_OUTFLOW_EQ
		|| g == GEO_OUTFLOW_RIGHT || g == GEO_NOTHING;
}

template <int STORE>
int step(const tnl_lbm_d3q27_data* SD, const void* lam_next, void* lam, double* grad)
{
	using S = Store<STORE>;
	using elem = typename S::elem;
	using N = Site27<STORE, TNL_LBM_COLL_CUM, TNL_LBM_PATTERN_AB>;
	const Geometry G{SD->X, SD->Y, SD->Z, SD->ox, SD->oy, SD->oz, SD->XYZ};
	const elem* F = (const elem*) SD->df[0];
	const elem* Ln = (const elem*) lam_next;
	elem* L = (elem*) lam;
	const int64_t NS = SD->X * SD->Y * SD->Z;
	std::vector<double> acc(27 * (size_t) SD->XYZ, 0.0);
	const bool dist[3] = {false, false, false};
	for (int64_t x = 0; x < SD->X; x++)
		for (int64_t z = 0; z < SD->Z; z++)
			for (int64_t y = 0; y < SD->Y; y++) {
				const int64_t s = G.s(x, y, z);
				const int16_t gi = SD->map ? SD->map[s] : (int16_t) GEO_PERIODIC;
				if (! supported(gi))
					return fail(TNL_LBM_ERR_UNSUPPORTED, "geometry code without an adjoint");
				if (gi == GEO_NOTHING)
					continue;
				int64_t xp, xm, yp, ym, zp, zm;
				N::indices(G, gi == GEO_PERIODIC, dist, x, y, z, xp, xm, yp, ym, zp, zm);
				int64_t xx = x;
				if (gi == GEO_OUTFLOW_RIGHT)
					xp = xx = xm;
				int64_t src[27];
				KS27<Dual> KS;
				for (int q = 0; q < 27; q++) {
					src[q] = G.s(N::nb(-CX[q], xm, xx, xp), N::nb(-CY[q], ym, y, yp), N::nb(-CZ[q], zm, z, zp));
					KS.f[q] = Dual((double) S::load(F, q * G.XYZ + src[q], q));
					KS.f[q].d[q] = 1;
				}
				const int64_t iy = SD->Y * z + y;
				Dual uin[3];
				uin[0] = Dual(inflow_at(SD, 0, iy));
				uin[1] = Dual(inflow_at(SD, 1, iy));
				uin[2] = Dual(inflow_at(SD, 2, iy));
				for (int d = 0; d < 3; d++)
					uin[d].d[27 + d] = 1;
				site_map(gi, KS, uin, SD->lbm_viscosity, SD->fx, SD->fy, SD->fz);
				double lam_out[27];
				for (int k = 0; k < 27; k++)
					lam_out[k] = (double) Ln[k * G.XYZ + s];
				for (int q = 0; q < 27; q++) {
					double m = 0;
					for (int k = 0; k < 27; k++)
						m += lam_out[k] * KS.f[k].d[q];
					acc[q * (size_t) G.XYZ + src[q]] += m;
				}
				if (grad)
					for (int d = 0; d < 3; d++) {
						double g = 0;
						for (int k = 0; k < 27; k++)
							g += lam_out[k] * KS.f[k].d[27 + d];
						grad[d * SD->Y * SD->Z + iy] += g;
					}
			}
	(void) NS;
	for (int q = 0; q < 27; q++)
		for (int64_t i = 0; i < G.XYZ; i++)
			L[q * G.XYZ + i] = (elem) acc[q * (size_t) G.XYZ + i];
	return TNL_LBM_OK;
}

}  // namespace adj

extern "C" {

const char* oracle_last_error(void)
{
	return g_err.c_str();
}

int oracle_d3q27_step(const tnl_lbm_config* cfg, const tnl_lbm_d3q27_data* SD, const int64_t offset[3],
					  const int64_t end[3], uint32_t flags, int nthreads)
{
	if (int rc = check(cfg, SD))
		return rc;
	STORE_SWITCH(cfg->storage, (dispatch_pattern<ST>(cfg, SD, offset, end, flags, nthreads)));
}

int oracle_nse_ade_step(const tnl_lbm_config* cfg, const tnl_lbm_d3q27_data* nse, const tnl_lbm_d3q7_data* ade,
						const int64_t offset[3], const int64_t end[3], uint32_t flags, int nthreads)
{
	if (int rc = check(cfg, nse))
		return rc;
	if (! ade || ade->struct_size != sizeof(tnl_lbm_d3q7_data) || ! ade->df[0])
		return fail(TNL_LBM_ERR_ARG, "bad ADE data");
	if (cfg->collision != TNL_LBM_COLL_CUM)
		return fail(TNL_LBM_ERR_UNSUPPORTED, "coupled oracle: default cumulant only");
#define CPL(ST)                                                                                                         \
	return cfg->pattern == TNL_LBM_PATTERN_AA                                                                           \
			 ? run_coupled<ST, TNL_LBM_COLL_CUM, TNL_LBM_PATTERN_AA>(nse, ade, offset, end, flags, cfg->macro_real, nthreads) \
			 : run_coupled<ST, TNL_LBM_COLL_CUM, TNL_LBM_PATTERN_AB>(nse, ade, offset, end, flags, cfg->macro_real, nthreads);
	switch (cfg->storage) {
		case TNL_LBM_STORE_F64: CPL(TNL_LBM_STORE_F64)
		case TNL_LBM_STORE_F32: CPL(TNL_LBM_STORE_F32)
		case TNL_LBM_STORE_F32_DEV: CPL(TNL_LBM_STORE_F32_DEV)
		case TNL_LBM_STORE_F16_DEV: CPL(TNL_LBM_STORE_F16_DEV)
		case TNL_LBM_STORE_BF16_DEV: CPL(TNL_LBM_STORE_BF16_DEV)
		default: CPL(TNL_LBM_STORE_I16_DEV)
	}
#undef CPL
}

int oracle_d3q27_set_equilibrium(const tnl_lbm_config* cfg, const tnl_lbm_d3q27_data* SD, double rho, double vx,
								 double vy, double vz)
{
	if (int rc = check(cfg, SD))
		return rc;
	STORE_SWITCH(cfg->storage, (set_eq_all<ST>(cfg, SD, nullptr, nullptr, nullptr, nullptr, rho, vx, vy, vz)));
}

int oracle_d3q27_set_equilibrium_field(const tnl_lbm_config* cfg, const tnl_lbm_d3q27_data* SD, const double* rho,
									   const double* vx, const double* vy, const double* vz)
{
	if (int rc = check(cfg, SD))
		return rc;
	if (! rho || ! vx || ! vy || ! vz)
		return fail(TNL_LBM_ERR_NULL, "field pointer missing");
	STORE_SWITCH(cfg->storage, (set_eq_all<ST>(cfg, SD, rho, vx, vy, vz, 0, 0, 0, 0)));
}

int oracle_d3q27_compute_macro(const tnl_lbm_config* cfg, const tnl_lbm_d3q27_data* SD)
{
	if (int rc = check(cfg, SD))
		return rc;
	if (! SD->macro)
		return fail(TNL_LBM_ERR_NULL, "macro array missing");
	STORE_SWITCH(cfg->storage, (compute_macro_all<ST>(cfg, SD)));
}

int oracle_d3q27_export_f64(const tnl_lbm_config* cfg, const tnl_lbm_d3q27_data* SD, int which, double* out)
{
	if (int rc = check(cfg, SD))
		return rc;
	if (which < 0 || which > 1 || ! SD->df[which])
		return fail(TNL_LBM_ERR_ARG, "bad array selector");
	STORE_SWITCH(cfg->storage, (export_all<ST>(SD, which, out, 27)));
}

int oracle_d3q27_import_f64(const tnl_lbm_config* cfg, const tnl_lbm_d3q27_data* SD, int which, const double* in)
{
	if (int rc = check(cfg, SD))
		return rc;
	if (which < 0 || which > 1 || ! SD->df[which])
		return fail(TNL_LBM_ERR_ARG, "bad array selector");
	STORE_SWITCH(cfg->storage, (import_all<ST>(SD, which, in)));
}

int oracle_d3q7_set_equilibrium(const tnl_lbm_config* cfg, const tnl_lbm_d3q7_data* ADE, const void* phi, double phi0)
{
	if (! cfg || ! ADE || ! ADE->df[0])
		return fail(TNL_LBM_ERR_NULL, "null");
	const int64_t NS = (ADE->X + 2 * ADE->ox) * (ADE->Y + 2 * ADE->oy) * (ADE->Z + 2 * ADE->oz);
	for (int64_t s = 0; s < NS; s++) {
		const double p = phi ? (cfg->macro_real == TNL_LBM_REAL_F64 ? ((const double*) phi)[s] : ((const float*) phi)[s]) : phi0;
		for (int q = 0; q < 7; q++) {
			const double v = Ade<double>::eq(q, p, 0, 0, 0);
			const float dv = (float) ((v - w7(q) * ADE->phi_ref) * ADE->store_scale);
			for (int k = 0; k < (cfg->pattern == TNL_LBM_PATTERN_AB ? 2 : 1); k++) {
				switch (cfg->storage) {
					case TNL_LBM_STORE_F64: ((double*) ADE->df[k])[q * ADE->XYZ + s] = v; break;
					case TNL_LBM_STORE_F32: ((float*) ADE->df[k])[q * ADE->XYZ + s] = (float) Ade<float>::eq(q, (float) p, 0, 0, 0); break;
					case TNL_LBM_STORE_F32_DEV: ((float*) ADE->df[k])[q * ADE->XYZ + s] = dv; break;
					case TNL_LBM_STORE_F16_DEV: ((uint16_t*) ADE->df[k])[q * ADE->XYZ + s] = f32_to_f16(dv); break;
					case TNL_LBM_STORE_BF16_DEV: ((uint16_t*) ADE->df[k])[q * ADE->XYZ + s] = f32_to_bf16(dv); break;
					default: ((int16_t*) ADE->df[k])[q * ADE->XYZ + s] = f32_to_snorm16(dv * (float) (1.0 / w7(q))); break;
				}
			}
		}
	}
	return TNL_LBM_OK;
}

int oracle_d3q7_export_f64(const tnl_lbm_config* cfg, const tnl_lbm_d3q7_data* ADE, int which, double* out)
{
	if (! cfg || ! ADE || which < 0 || which > 1 || ! ADE->df[which])
		return fail(TNL_LBM_ERR_NULL, "null");
	const int64_t NS = (ADE->X + 2 * ADE->ox) * (ADE->Y + 2 * ADE->oy) * (ADE->Z + 2 * ADE->oz);
	auto go = [&](auto stc) {
		constexpr int ST = decltype(stc)::value;
		using S7 = Store7<ST>;
		using real = typename S7::real;
		for (int q = 0; q < 7; q++)
			for (int64_t s = 0; s < NS; s++)
				out[q * ADE->XYZ + s] = S7::load((const typename S7::elem*) ADE->df[which], q * ADE->XYZ + s, q,
												 (real) ADE->phi_ref, (real) ADE->store_scale);
		return TNL_LBM_OK;
	};
	STORE_SWITCH(cfg->storage, go(std::integral_constant<int, ST>{}));
}

int oracle_d3q27_collide(int collision, int single, double* f, double nu, double fx, double fy, double fz, double* out)
{
	auto run = [&](auto zero) {
		using real = decltype(zero);
		KS27<real> KS;
		for (int q = 0; q < 27; q++)
			KS.f[q] = (real) f[q];
		KS.lbmViscosity = (real) nu;
		KS.fx = (real) fx;
		KS.fy = (real) fy;
		KS.fz = (real) fz;
		Abs<real>::densityVelocity(KS);
		if (out) {
			out[0] = KS.rho; out[1] = KS.vx; out[2] = KS.vy; out[3] = KS.vz;
		}
		Abs<real>::collide_any(KS, collision);
		for (int q = 0; q < 27; q++)
			f[q] = KS.f[q];
	};
	if (collision < TNL_LBM_COLL_CUM || collision > TNL_LBM_COLL_KBC_C4)
		return fail(TNL_LBM_ERR_UNSUPPORTED, "collision");
	if (single)
		run(0.0f);
	else
		run(0.0);
	return TNL_LBM_OK;
}

// deviation-form collision; `single` = 0 evaluates the same algebra in double (test hook)
int oracle_d3q27_collide_dev(int collision, double* g, double nu, double fx, double fy, double fz, int single)
{
	if (collision != TNL_LBM_COLL_CUM)
		return fail(TNL_LBM_ERR_UNSUPPORTED, "collision");
	auto run = [&](auto zero) {
		using real = decltype(zero);
		KS27<real> KS;
		for (int q = 0; q < 27; q++)
			KS.f[q] = (real) g[q];
		KS.lbmViscosity = (real) nu;
		KS.fx = (real) fx;
		KS.fy = (real) fy;
		KS.fz = (real) fz;
		DevT<real>::densityVelocity(KS);
		DevT<real>::cumulant(KS);
		for (int q = 0; q < 27; q++)
			g[q] = KS.f[q];
	};
	if (single)
		run(0.0f);
	else
		run(0.0);
	return TNL_LBM_OK;
}

int oracle_d3q27_equilibrium(int which_eq, int single, double rho, double vx, double vy, double vz, double* f)
{
	for (int q = 0; q < 27; q++)
		f[q] = single ? (double) Abs<float>::eq_any(which_eq, q, (float) rho, (float) vx, (float) vy, (float) vz)
					  : Abs<double>::eq_any(which_eq, q, rho, vx, vy, vz);
	return TNL_LBM_OK;
}

int oracle_d3q27_density_velocity(int single, const double* f, double fx, double fy, double fz, double* out4)
{
	auto run = [&](auto zero) {
		using real = decltype(zero);
		KS27<real> KS;
		for (int q = 0; q < 27; q++)
			KS.f[q] = (real) f[q];
		KS.fx = (real) fx;
		KS.fy = (real) fy;
		KS.fz = (real) fz;
		Abs<real>::densityVelocity(KS);
		out4[0] = KS.rho; out4[1] = KS.vx; out4[2] = KS.vy; out4[3] = KS.vz;
	};
	if (single)
		run(0.0f);
	else
		run(0.0);
	return TNL_LBM_OK;
}

uint16_t oracle_f32_to_f16(float v) { return f32_to_f16(v); }
float oracle_f16_to_f32(uint16_t h) { return f16_to_f32(h); }
uint16_t oracle_f32_to_bf16(float v) { return f32_to_bf16(v); }
float oracle_bf16_to_f32(uint16_t h) { return bf16_to_f32(h); }


// adjoint of one A-B step: lambda <- J^T lambda_next, grad[3][Y*Z] += d/d(inflow)
int oracle_d3q27_adjoint_step(const tnl_lbm_config* cfg, const tnl_lbm_d3q27_data* SD, const void* lambda_next,
							  void* lambda, double* grad_inflow)
{
	if (int rc = check(cfg, SD))
		return rc;
	if (cfg->collision != TNL_LBM_COLL_CUM || cfg->pattern != TNL_LBM_PATTERN_AB)
		return fail(TNL_LBM_ERR_UNSUPPORTED, "the adjoint is built for the default cumulant, A-B pattern");
	if (cfg->storage != TNL_LBM_STORE_F64 && cfg->storage != TNL_LBM_STORE_F32)
		return fail(TNL_LBM_ERR_UNSUPPORTED, "the adjoint is built for the absolute storage forms");
	if (SD->ox || SD->oy || SD->oz)
		return fail(TNL_LBM_ERR_UNSUPPORTED, "the adjoint is built for a single block without overlaps");
	if (SD->force_field)
		return fail(TNL_LBM_ERR_UNSUPPORTED, "the adjoint is built for the homogeneous force only");
	if (! lambda_next || ! lambda)
		return fail(TNL_LBM_ERR_NULL, "adjoint arrays missing");
	if (cfg->storage == TNL_LBM_STORE_F64)
		return adj::step<TNL_LBM_STORE_F64>(SD, lambda_next, lambda, grad_inflow);
	return adj::step<TNL_LBM_STORE_F32>(SD, lambda_next, lambda, grad_inflow);
}

}  // extern "C"
