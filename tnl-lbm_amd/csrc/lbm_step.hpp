// Device side of the D3Q27 step shared by the library's translation units:
// kernel arguments, the periodic A-A fast kernel, the general (map + BC)
// kernel, the coupled NSE + ADE kernel, initialisation / transfer /
// reduction kernels, and the launch template.  See lbm_d3q27.hip for the
// design notes and the C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

#include "lbm_ade.hpp"
#include "lbm_device.hpp"

namespace lbm {

// host-side error reporting (defined in lbm_d3q27.hip)
int hip_check(hipError_t e, const char* what);

// Whether the kernels that can address a population array of at most 4 GiB
// as a per-direction base plus a 32-bit offset may do so; false while
// tnl_lbm_set_wide_addressing(1) holds (defined in lbm_d3q27.hip).
bool narrow_slots_allowed();

// a population array of XYZ elements of `elem` bytes takes 32-bit slot offsets
inline bool narrow_slots(int64_t XYZ, size_t elem)
{
	return narrow_slots_allowed() && (uint64_t) XYZ * elem <= (1ull << 32);
}

// ---------------------------------------------------------------------------
struct StepArgs
{
	void* df0;
	void* df1;
	void* macro;
	const int16_t* map;
	int64_t XYZ;
	int64_t X, Y, Z;     // sizes without overlaps
	int64_t ox, oy, oz;  // overlaps
	int64_t Ys, Zs;      // stored extents Y+2oy, Z+2oz
	uint64_t plane_bytes;  // XYZ * sizeof(element): byte stride between population planes
	float rs[4], invr[4];  // fixed-point form per direction class: R_n / 32767 and 1 / R_n
	// fixed-point form per weight class n = |c_q|^2 (w_n = (2/3)^(3-n) / 6^n):
	// decode w_n R / 32767, encode 1 / (w_n R) times the fast collision's
	// 2^-n output factor
	float cls_sc[4], cls_inv[4];
	int32_t offx, offy, offz;
	int32_t nx, ny, nz;
	uint32_t nyb;      // y segments per row (linear mapping: blocks per x plane)
	uint32_t nblocks;  // logical blocks
	uint32_t per_xcd;  // padded blocks / 8
	int32_t lin;       // linear y-z mapping (see linearize)
	int32_t even;
	int32_t out_macro;
	int32_t macro_f64;
	int32_t dist_x, dist_y, dist_z;
	double omega1;
	double nu;  // lbmViscosity (MRT_LES and KBC use it directly)
	double fx, fy, fz;
	double inflow[3];
	const void* prof[3];  // inflow profiles: double arrays, or float where prof_f32
	int32_t prof_f32;
	const void* force;  // per-site force: 3 planes of macro elements (macro_f64), or null
	void* tape;  // moment recording of the A-B step (10 planes of XYZ), or null
	// a launch over one row list (tnl_lbm_classify_rows): the listed (x, z)
	// rows of the box, each entry x << 16 | z; null = the box itself
	const uint32_t* rows;
	uint32_t nrows;
	// host side: the three lists of a full-box step (classes 0, 1, 2 one
	// after another), or null
	const uint32_t* lists;
	int64_t list_count[3];
};

constexpr int NXCD = 8;

// per-site force of MACRO::computeForcing (kernels.h:103): added to the
// homogeneous force of copyQuantities (macro.h:80-86) in the arithmetic of
// the collision
template <typename real>
LBM_DEV void add_site_force(const StepArgs& A, int64_t s, real& Fx, real& Fy, real& Fz)
{
	if (A.macro_f64) {
		const double* F = (const double*) A.force;
		Fx += (real) F[s];
		Fy += (real) F[A.XYZ + s];
		Fz += (real) F[2 * A.XYZ + s];
	}
	else {
		const float* F = (const float*) A.force;
		Fx += (real) F[s];
		Fy += (real) F[A.XYZ + s];
		Fz += (real) F[2 * A.XYZ + s];
	}
}

// component k of the inflow velocity at profile index i = Y z + y
// (NSE_Data_InflowProfile::vel_*, lbm_data.h:104-125; the constant of
// NSE_Data_ConstInflow where no profile is given)
LBM_DEV double inflow_component(const StepArgs& A, int k, int64_t i)
{
	if (! A.prof[k])
		return A.inflow[k];
	return A.prof_f32 ? (double) static_cast<const float*>(A.prof[k])[i] : static_cast<const double*>(A.prof[k])[i];
}

// logical block -> (x, z, y segment) with XCD-aware remapping
// LIN = false: whole rows only (x and z stay uniform per block, as the
// periodic kernel's scalar addressing needs)
// LIST: a launch over a row list (A.rows); a compile-time choice, since
// coordinates that may come from either source are per-lane values in both
// (the coupled kernel's registers rose from 109 to 133 with a run-time one)
template <bool LIN = true, bool LIST = false>
LBM_DEV bool block_coords(const StepArgs& A, int64_t& x, int64_t& z, int64_t& y)
{
	const uint32_t b = blockIdx.x;
	const uint32_t L = (b % NXCD) * A.per_xcd + b / NXCD;
	if (L >= A.nblocks)
		return false;
	if constexpr (LIST) {
		// a row list: whole rows per block, or the listed rows' sites one
		// after another (linear mapping)
		uint32_t k, yy;
		if (LIN && A.lin) {
			const uint32_t i = L * blockDim.x + threadIdx.x;
			k = i / (uint32_t) A.ny;
			if (k >= A.nrows)
				return false;
			yy = i % (uint32_t) A.ny;
		}
		else {
			k = L / A.nyb;
			yy = (L % A.nyb) * blockDim.x + threadIdx.x;
			if (yy >= (uint32_t) A.ny)
				return false;
		}
		const uint32_t r = A.rows[k];
		x = (int64_t) (r >> 16);
		z = (int64_t) (r & 0xffffu);
		y = (int64_t) yy;
		return x < A.X && z < A.Z;
	}
	if (LIN && A.lin) {
		// consecutive lanes walk the y-z plane of one x, row after row
		const uint32_t i = (L % A.nyb) * blockDim.x + threadIdx.x;
		x = A.offx + (int64_t) (L / A.nyb);
		if (i >= (uint32_t) A.ny * (uint32_t) A.nz)
			return false;
		z = A.offz + (int64_t) (i / (uint32_t) A.ny);
		y = A.offy + (int64_t) (i % (uint32_t) A.ny);
		return true;
	}
	const uint32_t yb = L % A.nyb;
	const uint32_t t = L / A.nyb;
	z = A.offz + (int64_t) (t % (uint32_t) A.nz);
	x = A.offx + (int64_t) (t / (uint32_t) A.nz);
	y = A.offy + (int64_t) yb * blockDim.x + threadIdx.x;
	return y < A.offy + A.ny;
}

template <typename E>
LBM_DEV E* at(E* base, uint32_t i)
{
	// uniform 64-bit base + 32-bit per-lane byte offset (global saddr form)
	return (E*) ((char*) base + (uint32_t) (i * (uint32_t) sizeof(E)));
}

template <typename real>
LBM_DEV void write_macro(const StepArgs& A, int64_t s, real rho, real vx, real vy, real vz)
{
	if (A.macro_f64) {
		double* m = (double*) A.macro;
		m[s] = rho;
		m[A.XYZ + s] = vx;
		m[2 * A.XYZ + s] = vy;
		m[3 * A.XYZ + s] = vz;
	}
	else {
		float* m = (float*) A.macro;
		m[s] = (float) rho;
		m[A.XYZ + s] = (float) vx;
		m[2 * A.XYZ + s] = (float) vy;
		m[3 * A.XYZ + s] = (float) vz;
	}
}

// COLL::collision after computeDensityAndVelocity, in the reference's operation order
template <int ST, int COLL>
LBM_DEV void collide(typename Storage<ST>::real* f, typename Storage<ST>::real rho, typename Storage<ST>::real vx,
					 typename Storage<ST>::real vy, typename Storage<ST>::real vz, const StepArgs& A,
					 typename Storage<ST>::real Fx, typename Storage<ST>::real Fy, typename Storage<ST>::real Fz)
{
	using P = Phys<ST>;
	using real = typename Storage<ST>::real;
	const real omega1 = (real) A.omega1;
	if constexpr (Storage<ST>::dev)
		P::cumulant(f, rho, vx, vy, vz, omega1);
	else if constexpr (COLL == TNL_LBM_COLL_CUM)
		P::template cumulant<false, false>(f, rho, vx, vy, vz, omega1);
	else if constexpr (COLL == TNL_LBM_COLL_CUM_2017)
		P::template cumulant<true, false>(f, rho, vx, vy, vz, omega1);
	else if constexpr (COLL == TNL_LBM_COLL_CUM_2017_ANTIALIAS)
		P::template cumulant<true, true>(f, rho, vx, vy, vz, omega1);
	else if constexpr (COLL == TNL_LBM_COLL_MRT_LES)
		P::mrt(f, rho, vx, vy, vz, (real) A.nu);
	else {
		// KBC_N1..N4 = 4..7, KBC_C1..C4 = 8..11: shear part D, D+T, D+Q, D+T+Q
		constexpr int k = (COLL - TNL_LBM_COLL_KBC_N1) % 4;
		constexpr bool central = COLL >= TNL_LBM_COLL_KBC_C1;
		P::template kbc<central, k == 1 || k == 3, k >= 2>(f, rho, vx, vy, vz, (real) A.nu, Fx, Fy, Fz);
	}
}

// ===========================================================================
// Fast path: A-A pattern, every site GEO_PERIODIC (no map).
// Even step (streaming_AA.h:81-84, 16-45): read own slots, write opposite.
// Odd step (streaming_AA.h:86-115, 46-75): pull from r - c_q (opposite
// slot), push to r + c_q (same slot).
// Addresses: base + plane[q] (kernel argument, scalar) + row (scalar) + y
// (per lane), one vector add per access.  The even step streams with
// nontemporal hints; so do the odd step's accesses in the site's own y,
// while its +-1 shifted rows share cache lines between neighbouring waves
// and keep the default policy.
// ===========================================================================
template <int ST, int COLL>
constexpr bool fast_collision = std::is_same<typename Storage<ST>::real, float>::value && COLL == TNL_LBM_COLL_CUM;

// Collision slot of the streaming-only measurement kernel
// (tnl_lbm_d3q27_stream_only): the identity, so that the periodic kernels
// move the populations of free streaming with the step's own loads, stores
// and addressing, and time the memory traffic of the access pattern alone.
constexpr int COLL_STREAM_ONLY = 1000;

// number of non-zero components of c_q (the weight class of q)
constexpr int ncls(int q)
{
	return (CX[q] != 0) + (CY[q] != 0) + (CZ[q] != 0);
}

// Factor applied to a post-collision value before its (non-fixed-point)
// store: the fast collision leaves population q scaled by 2^n(q).
template <int ST, int COLL>
constexpr float out_factor(int q)
{
	return fast_collision<ST, COLL> ? FastCum<Storage<ST>::dev>::unscale(q) : 1.0f;
}

template <int ST, int COLL>
LBM_DEV void collide_site(typename Storage<ST>::real* f, const StepArgs& A, typename Storage<ST>::real& rho,
						  typename Storage<ST>::real& vx, typename Storage<ST>::real& vy, typename Storage<ST>::real& vz)
{
	using real = typename Storage<ST>::real;
	if constexpr (COLL == COLL_STREAM_ONLY) {
		rho = vx = vy = vz = 0;
	}
	else if constexpr (fast_collision<ST, COLL>) {
		FastCum<Storage<ST>::dev>::collide(f, (real) A.fx, (real) A.fy, (real) A.fz, (real) A.omega1, true, rho, vx, vy, vz);
	}
	else {
		Phys<ST>::density_velocity(f, (real) A.fx, (real) A.fy, (real) A.fz, rho, vx, vy, vz);
		collide<ST, COLL>(f, rho, vx, vy, vz, A, (real) A.fx, (real) A.fy, (real) A.fz);
	}
}

// Keeps `p` in scalar registers and makes everything derived from it depend
// on `v`: the store addresses are then formed after the collision instead of
// being hoisted above it (which would keep 27 scalar pairs live and spill).
template <bool PIN, typename T>
LBM_DEV char* after(char* p, const T& v)
{
	if constexpr (PIN) {
		auto g = (__attribute__((address_space(1))) char*) p;
		asm volatile("" : "+s"(g) : "v"(v));
		return (char*) g;
	}
	return p;
}

// Build-time tuning knobs (defaults are the measured best on MI355X):
//  LBM_NT_EVEN   bitmask over storage forms: nontemporal hints on the even step
//  LBM_PIN_ADDR  bitmask over storage forms: scalar base pointers, even step
//  LBM_NT_ODD    bitmask over storage forms: nontemporal hints on the odd
//                step's accesses within the site's own y (c_y = 0); the +-1
//                shifted rows share lines with the neighbouring wave through
//                L2 and keep the default policy (with hints they measured 40%
//                slower)
//  LBM_PIN_ODD   bitmask over storage forms: scalar base pointers, odd step
//                (F16 / BF16 odd step at 512^3, one allocation: -3.8% / -1.3%
//                time; HBM writes above the algorithmic bytes +2.8% -> +0.8% /
//                +5.2% -> +1.7%)
//  LBM_WAVES_F64 minimum waves per SIMD requested for the FP64 kernels
#ifndef LBM_NT_EVEN
#define LBM_NT_EVEN 0x3f
#endif
#ifndef LBM_PIN_ADDR
#define LBM_PIN_ADDR 0x3f
#endif
#ifndef LBM_NT_ODD
#define LBM_NT_ODD 0x3f
#endif
//  LBM_NT_GENERAL bitmask over storage forms: the same hints in the map kernel
#ifndef LBM_NT_GENERAL
#define LBM_NT_GENERAL 0x3f
#endif
#ifndef LBM_PIN_ODD
#define LBM_PIN_ODD 0x3f
#endif
//  LBM_TPB_FP32  block size of the periodic kernel for the FP32-compute forms
//                where the row length is a multiple of it (the paired 2-byte
//                even step keeps its own shape): a 512-site row in one block
//                keeps 7 of its 8 wave boundaries, whose +-1 shifted lines the
//                neighbouring waves share on the odd step, on one CU
//                (measured +0.6 to +1.2 points on both steps at 512^3)
#ifndef LBM_TPB_FP32
#define LBM_TPB_FP32 512
#endif
#ifndef LBM_WAVES_F64
#define LBM_WAVES_F64 1
#endif
//  LBM_PAIR_EVEN bitmask over the 2-byte storage forms: two sites per lane
//                on the even step (k_aa_even_pair)
#ifndef LBM_PAIR_EVEN
#define LBM_PAIR_EVEN 0x38
#endif
#ifndef LBM_MAXWAVES_F64
#define LBM_MAXWAVES_F64 3
#endif

// Pins a wave-uniform pointer (base + plane + row) in scalar registers, so
// that each access is one global load/store with a scalar base and a 32-bit
// per-lane offset; left to itself the compiler adds the shared per-lane part
// first and then pays a 64-bit vector add per access.
// (The pinned copy is a global-address-space pointer: a plain pointer that
// has passed through the asm would make every access a flat one.)
template <bool PIN>
LBM_DEV char* sptr(char* p)
{
	if constexpr (PIN) {
		auto g = (__attribute__((address_space(1))) char*) p;
		asm volatile("" : "+s"(g));
		return (char*) g;
	}
	return p;
}

template <int ST, int COLL, bool EVEN>
__global__ void __launch_bounds__(ST == TNL_LBM_STORE_F64 ? 256 : LBM_TPB_FP32)
__attribute__((amdgpu_waves_per_eu(ST == TNL_LBM_STORE_F64 ? LBM_WAVES_F64 : 1, ST == TNL_LBM_STORE_F64 ? LBM_MAXWAVES_F64 : 8)))
k_aa_periodic(const StepArgs A)
{
	using S = Storage<ST>;
	using real = typename S::real;
	using elem = typename S::elem;
	constexpr uint32_t E = sizeof(elem);
	constexpr bool NT = (LBM_NT_EVEN >> ST) & 1;
	constexpr bool PIN = (LBM_PIN_ADDR >> ST) & 1;
	constexpr bool PIN_ODD = (LBM_PIN_ODD >> ST) & 1;
	constexpr bool NT_ODD = (LBM_NT_ODD >> ST) & 1;
	int64_t x, z, y;
	if (! block_coords<false>(A, x, z, y))
		return;
	char* __restrict__ df = (char*) A.df0;
	const uint64_t pb = A.plane_bytes;
	real f[27];
	const uint32_t yl = (uint32_t) (y + A.oy);
	if constexpr (EVEN) {
		const uint64_t row = (uint64_t) (((x + A.ox) * A.Zs + (z + A.oz)) * A.Ys) * E;
		const uint32_t yb = yl * E;
#pragma unroll
		for (int q = 0; q < 27; q++) {
			const elem* p = (const elem*) (sptr<PIN>(df + ((uint64_t) q * pb + row)) + yb);
			const elem v = NT ? ld_nt(p) : *p;
			f[q] = S::ld(&v, A.cls_sc[ncls(q)]);
		}
		real rho, vx, vy, vz;
		collide_site<ST, COLL>(f, A, rho, vx, vy, vz);
		char* out = after<PIN>(df, f[26]);
		if constexpr (S::fixed) {
			// two populations per conversion
#pragma unroll
			for (int q = 0; q < 27; q += 2) {
				const int q1 = q + 1 < 27 ? q + 1 : q;
				const auto r = __builtin_amdgcn_cvt_pknorm_i16(f[q] * A.cls_inv[ncls(q)], f[q1] * A.cls_inv[ncls(q1)]);
				elem* p0 = (elem*) (sptr<PIN>(out + ((uint64_t) opp(q) * pb + row)) + yb);
				NT ? st_nt(p0, (elem) r[0]) : (void) (*p0 = r[0]);
				if (q + 1 < 27) {
					elem* p1 = (elem*) (sptr<PIN>(out + ((uint64_t) opp(q1) * pb + row)) + yb);
					NT ? st_nt(p1, (elem) r[1]) : (void) (*p1 = r[1]);
				}
			}
		}
		else {
#pragma unroll
			for (int q = 0; q < 27; q++) {
				elem v;
				S::st(&v, f[q] * out_factor<ST, COLL>(q));
				elem* p = (elem*) (sptr<PIN>(out + ((uint64_t) opp(q) * pb + row)) + yb);
				if (NT)
					st_nt(p, v);
				else
					*p = v;
			}
		}
		if (A.out_macro)
			write_macro(A, (int64_t) (row / E) + yl, rho, vx, vy, vz);
	}
	else {
		// periodic neighbours (kernelInitIndices, kernels.h:21-28)
		const int64_t xp = (! A.dist_x && x == A.X - 1) ? 0 : x + 1;
		const int64_t xm = (! A.dist_x && x == 0) ? A.X - 1 : x - 1;
		const int64_t zp = (! A.dist_z && z == A.Z - 1) ? 0 : z + 1;
		const int64_t zm = (! A.dist_z && z == 0) ? A.Z - 1 : z - 1;
		const uint32_t ypl = (uint32_t) (((! A.dist_y && y == A.Y - 1) ? 0 : y + 1) + A.oy);
		const uint32_t yml = (uint32_t) (((! A.dist_y && y == 0) ? A.Y - 1 : y - 1) + A.oy);
		const int64_t xs[3] = {xm + A.ox, x + A.ox, xp + A.ox};
		const int64_t zs[3] = {zm + A.oz, z + A.oz, zp + A.oz};
		const uint32_t ys[3] = {yml * E, yl * E, ypl * E};
#pragma unroll
		for (int q = 0; q < 27; q++) {
			const uint64_t row = (uint64_t) ((xs[1 - CX[q]] * A.Zs + zs[1 - CZ[q]]) * A.Ys) * E;
			const elem* p = (const elem*) (sptr<PIN_ODD>(df + ((uint64_t) opp(q) * pb + row)) + ys[1 - CY[q]]);
			if (NT_ODD && CY[q] == 0) {
				const elem v = ld_nt(p);
				f[q] = S::ld(&v, A.cls_sc[ncls(q)]);
			}
			else
				f[q] = S::ld(p, A.cls_sc[ncls(q)]);
		}
		real rho, vx, vy, vz;
		collide_site<ST, COLL>(f, A, rho, vx, vy, vz);
		char* out = after<PIN_ODD>(df, f[26]);
		auto dst = [&](int q) {
			const uint64_t row = (uint64_t) ((xs[1 + CX[q]] * A.Zs + zs[1 + CZ[q]]) * A.Ys) * E;
			return (elem*) (sptr<PIN_ODD>(out + ((uint64_t) q * pb + row)) + ys[1 + CY[q]]);
		};
		auto put = [&](int q, elem v) {
			if (NT_ODD && CY[q] == 0)
				st_nt(dst(q), v);
			else
				*dst(q) = v;
		};
		if constexpr (S::fixed) {
			// two populations per conversion
#pragma unroll
			for (int q = 0; q < 27; q += 2) {
				const int q1 = q + 1 < 27 ? q + 1 : q;
				const auto r = __builtin_amdgcn_cvt_pknorm_i16(f[q] * A.cls_inv[ncls(q)], f[q1] * A.cls_inv[ncls(q1)]);
				put(q, (elem) r[0]);
				if (q + 1 < 27)
					put(q1, (elem) r[1]);
			}
		}
		else {
#pragma unroll
			for (int q = 0; q < 27; q++) {
				elem v;
				S::st(&v, f[q] * out_factor<ST, COLL>(q));
				put(q, v);
			}
		}
		if (A.out_macro)
			write_macro(A, (xs[1] * A.Zs + zs[1]) * A.Ys + yl, rho, vx, vy, vz);
	}
}

// ===========================================================================
// Paired even step for the 2-byte storage forms (the A-A even step above).
// Each lane updates the two sites y, y + 1 of a row (y even) and moves both
// through one aligned 4-byte access per population, site y in the low half:
// half the memory instructions per site.  The host uses it where Y, oy, offy,
// ny and XYZ are even.  (The odd step keeps one site per lane: its +-1
// shifted pairs would need 2-byte aligned 4-byte accesses, which measured
// slower than the single 2-byte accesses they replace.)
// ===========================================================================
template <int ST, int COLL>
struct PairCodec
{
	using S = Storage<ST>;
	using elem = typename S::elem;
	static_assert(sizeof(elem) == 2, "pairs of 2-byte elements");
	static LBM_DEV float get(uint32_t h, int q, const StepArgs& A)
	{
		if constexpr (S::fixed)
			return (float) (int16_t) (uint16_t) h * A.cls_sc[ncls(q)];
		else
			return (float) __builtin_bit_cast(elem, (uint16_t) h);
	}
	static LBM_DEV uint32_t put(float v, int q, const StepArgs& A)
	{
		if constexpr (S::fixed)
			return (uint16_t) __builtin_amdgcn_cvt_pknorm_i16(v * A.cls_inv[ncls(q)], 0.0f)[0];
		else
			return __builtin_bit_cast(uint16_t, (elem) (v * out_factor<ST, COLL>(q)));
	}
	// forms whose conversion packs two values at once
	static constexpr bool PACK2 = S::fixed || std::is_same<elem, __bf16>::value;
	// the value the conversion takes, and the packed pair (lo = site y)
	static LBM_DEV float scaled(float v, int q, const StepArgs& A)
	{
		if constexpr (S::fixed)
			return v * A.cls_inv[ncls(q)];
		else
			return v * out_factor<ST, COLL>(q);
	}
	static LBM_DEV uint32_t pack2(float lo, float hi)
	{
		if constexpr (S::fixed)
			return __builtin_bit_cast(uint32_t, __builtin_amdgcn_cvt_pknorm_i16(lo, hi));
		else {
			typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
			const bf16x2 v = {(__bf16) lo, (__bf16) hi};
			return __builtin_bit_cast(uint32_t, v);
		}
	}
};

// Collides the pair held in w[] in place: site s (low halves), then s + 1.
// For the float forms site s's outputs replace the low halves while the
// high halves still hold site s + 1's inputs, so one packed array serves
// both sites (binary16 writes a half and keeps the other in one
// instruction).
template <int ST, int COLL>
LBM_DEV void collide_pair(uint32_t* w, const StepArgs& A, int64_t s)
{
	using C = PairCodec<ST, COLL>;
	float f[27];
	float rho, vx, vy, vz;
	if constexpr (C::PACK2) {
		// I16 and BF16 convert two values per instruction: site s's scaled
		// outputs wait as floats, in the registers its decoded inputs held,
		// and one conversion packs both sites, with no per-half merges
		// (I16: 1001 -> 947 VALU per pair).
		float a[27];
#pragma unroll
		for (int q = 0; q < 27; q++)
			f[q] = C::get(w[q] & 0xffffu, q, A);
		collide_site<ST, COLL>(f, A, rho, vx, vy, vz);
		if (A.out_macro)
			write_macro(A, s, rho, vx, vy, vz);
#pragma unroll
		for (int q = 0; q < 27; q++)
			a[q] = C::scaled(f[q], q, A);
#pragma unroll
		for (int q = 0; q < 27; q++)
			f[q] = C::get(w[q] >> 16, q, A);
		collide_site<ST, COLL>(f, A, rho, vx, vy, vz);
		if (A.out_macro)
			write_macro(A, s + 1, rho, vx, vy, vz);
#pragma unroll
		for (int q = 0; q < 27; q++)
			w[q] = C::pack2(a[q], C::scaled(f[q], q, A));
		return;
	}
#pragma unroll
	for (int q = 0; q < 27; q++)
		f[q] = C::get(w[q] & 0xffffu, q, A);
	collide_site<ST, COLL>(f, A, rho, vx, vy, vz);
	if (A.out_macro)
		write_macro(A, s, rho, vx, vy, vz);
#pragma unroll
	for (int q = 0; q < 27; q++)
		w[q] = (w[q] & 0xffff0000u) | C::put(f[q], q, A);
#pragma unroll
	for (int q = 0; q < 27; q++)
		f[q] = C::get(w[q] >> 16, q, A);
	collide_site<ST, COLL>(f, A, rho, vx, vy, vz);
	if (A.out_macro)
		write_macro(A, s + 1, rho, vx, vy, vz);
#pragma unroll
	for (int q = 0; q < 27; q++)
		w[q] = (w[q] & 0xffffu) | (C::put(f[q], q, A) << 16);
}

template <int ST, int COLL>
__global__ void __launch_bounds__(256) k_aa_even_pair(const StepArgs A)
{
	constexpr bool NT = (LBM_NT_EVEN >> ST) & 1;
	constexpr bool PIN = (LBM_PIN_ADDR >> ST) & 1;
	const uint32_t b = blockIdx.x;
	const uint32_t L = (b % NXCD) * A.per_xcd + b / NXCD;
	if (L >= A.nblocks)
		return;
	const uint32_t t = L / A.nyb;
	const int64_t z = A.offz + (int64_t) (t % (uint32_t) A.nz);
	const int64_t x = A.offx + (int64_t) (t / (uint32_t) A.nz);
	const uint32_t pair = (L % A.nyb) * blockDim.x + threadIdx.x;
	if (2 * pair >= (uint32_t) A.ny)
		return;
	const uint32_t yl = (uint32_t) (A.offy + 2 * (int64_t) pair + A.oy);  // sites yl, yl + 1 of the stored row
	char* __restrict__ df = (char*) A.df0;
	const uint64_t pb = A.plane_bytes;
	const uint64_t row = (uint64_t) (((x + A.ox) * A.Zs + (z + A.oz)) * A.Ys) * 2;
	const uint32_t yb = yl * 2;
	uint32_t w[27];
#pragma unroll
	for (int q = 0; q < 27; q++) {
		const uint32_t* p = (const uint32_t*) (sptr<PIN>(df + ((uint64_t) q * pb + row)) + yb);
		w[q] = NT ? __builtin_nontemporal_load(p) : *p;
	}
	collide_pair<ST, COLL>(w, A, (int64_t) (row / 2) + yl);
	// The stores use the loads' lane offset, kept opaque: when the compiler
	// sees the two equal it shares 64-bit lane pointers between loads and
	// stores and drops both out of the scalar-base (saddr) form.  Each store's
	// base is pinned after its value is final (one base after the last value
	// left 64-bit vector adds on every store).
	uint32_t ob = yb;
	asm volatile("" : "+v"(ob));
#pragma unroll
	for (int q = 0; q < 27; q++) {
		uint32_t* p = (uint32_t*) ((char*) sptr<PIN>(after<PIN>(df, w[q]) + ((uint64_t) opp(q) * pb + row)) + ob);
		if (NT)
			__builtin_nontemporal_store(w[q], p);
		else
			*p = w[q];
	}
}

// ===========================================================================
// General path: geometry map + D3Q27_BC_All, A-A or A-B.
// ===========================================================================
enum : int16_t
{
	GEO_FLUID = 0, GEO_WALL, GEO_INFLOW, GEO_INFLOW_LEFT, GEO_INFLOW_BB_LEFT, GEO_INFLOW_EQ_LEFT, GEO_OUTFLOW_EQ,
	GEO_OUTFLOW_RIGHT, GEO_OUTFLOW_RIGHT_INTERP, GEO_PERIODIC, GEO_NOTHING, GEO_SYM_TOP, GEO_SYM_BOTTOM, GEO_SYM_LEFT,
	GEO_SYM_RIGHT, GEO_SYM_BACK, GEO_SYM_FRONT
};

struct Nb
{
	int64_t m, o, p;  // storage coordinates (overlap-shifted) of c = -1, 0, +1
	LBM_DEV int64_t operator[](int c) const { return c < 0 ? m : (c > 0 ? p : o); }
};

// kernelInitIndices (kernels.h:5-57).  A-A indices of non-periodic sites
// that would leave the stored array wrap instead (the reference reads
// outside its allocation there).
template <bool AA>
LBM_DEV Nb neighbours(int64_t c, int64_t n, int64_t o, bool dist, bool periodic)
{
	int64_t p, m;
	if (periodic) {
		p = (! dist && c == n - 1) ? 0 : c + 1;
		m = (! dist && c == 0) ? n - 1 : c - 1;
	}
	else if (AA) {
		p = c + 1 < n + o ? c + 1 : 0;
		m = c - 1 >= -o ? c - 1 : n - 1;
	}
	else {
		p = c + 1 < n - 1 + o ? c + 1 : n - 1 + o;
		m = c - 1 > -o ? c - 1 : -o;
	}
	return Nb{m + o, c + o, p + o};
}

// Moment recording layout: blocks of 64 consecutive sites, each holding its
// 10 moments as 10 runs of 64 values, so that one wave writes (and the
// adjoint reads) one contiguous 2560-element stretch instead of 10 planes.
__host__ __device__ inline int64_t tape_index(int k, int64_t s)
{
	return ((s >> 6) * 10 + k) * 64 + (s & 63);
}

// The raw moments m0, j, Pi of 27 populations, in one summation order shared
// by the step's moment recording and the adjoint that reads it back.
template <typename real>
LBM_DEV void raw_moments10(const real* f, real* m)
{
	real m0 = 0, jx = 0, jy = 0, jz = 0, Pxx = 0, Pyy = 0, Pzz = 0, Pxy = 0, Pxz = 0, Pyz = 0;
#pragma unroll
	for (int q = 0; q < 27; q++) {
		const real v = f[q];
		m0 += v;
		if (CX[q]) { jx += CX[q] > 0 ? v : -v; Pxx += v; }
		if (CY[q]) { jy += CY[q] > 0 ? v : -v; Pyy += v; }
		if (CZ[q]) { jz += CZ[q] > 0 ? v : -v; Pzz += v; }
		if (CX[q] * CY[q]) Pxy += CX[q] * CY[q] > 0 ? v : -v;
		if (CX[q] * CZ[q]) Pxz += CX[q] * CZ[q] > 0 ? v : -v;
		if (CY[q] * CZ[q]) Pyz += CY[q] * CZ[q] > 0 ? v : -v;
	}
	m[0] = m0; m[1] = jx; m[2] = jy; m[3] = jz; m[4] = Pxx; m[5] = Pyy; m[6] = Pzz; m[7] = Pxy; m[8] = Pxz; m[9] = Pyz;
}

// Slot q of the site at storage coordinates (xs, ys, zs), and of site s:
// with O32, the array of direction q as a wave-uniform base plus a 32-bit
// per-lane byte offset formed in 32-bit arithmetic; else a 64-bit address
// per lane. PIN keeps the base in scalar registers (one scalar-base access
// each); it needs the call to sit where the compiler can prove the base
// uniform, which the adjoint's branchy body does not allow.
template <bool O32, bool PIN = true, typename E>
LBM_DEV E* slot(E* base, int q, const StepArgs& A, int64_t xs, int64_t ys, int64_t zs)
{
	if constexpr (O32) {
		const uint32_t i = ((uint32_t) xs * (uint32_t) A.Zs + (uint32_t) zs) * (uint32_t) A.Ys + (uint32_t) ys;
		return at((E*) sptr<PIN>((char*) (base + q * A.XYZ)), i);
	}
	else
		return base + q * A.XYZ + (xs * A.Zs + zs) * A.Ys + ys;
}
template <bool O32, bool PIN = true, typename E>
LBM_DEV E* slot0(E* base, int q, const StepArgs& A, int64_t s)
{
	if constexpr (O32)
		return at((E*) sptr<PIN>((char*) (base + q * A.XYZ)), (uint32_t) s);
	else
		return base + q * A.XYZ + s;
}

// O32: every population array holds under 4 GiB (XYZ * element size <= 2^32,
// checked on the host), so a population slot is a wave-uniform base (the
// array of direction q, pinned in scalar registers) plus a 32-bit per-lane
// byte offset. The site index is then formed in 32-bit arithmetic and each
// access is one scalar-base load or store, in place of a 64-bit address per
// lane and access.
//
// LEAN: the sites of the launch carry only FLUID or PERIODIC codes (1), or
// those and WALL (2), as the row lists of tnl_lbm_classify_rows guarantee;
// every other boundary path is left out of the kernel.  (The kernels take
// LEAN 3 for the general paths over a row list: General sees 0.)  A kernel's register
// allocation is the maximum over all of its paths, so the rarely taken ones
// set it for every site: the F16 A-B map kernel needs 97 VGPRs with all of
// them, 49 with FLUID and PERIODIC alone.
template <int ST, int COLL, int PATTERN, int TAPE = 0, bool O32 = false, int LEAN = 0>
struct General
{
	using S = Storage<ST>;
	using real = typename S::real;
	using elem = typename S::elem;
	using P = Phys<ST>;
	static constexpr bool AA = PATTERN == TNL_LBM_PATTERN_AA;
	static constexpr int EQK = eq_kind_of(COLL);
	static constexpr bool FAST = fast_collision<ST, COLL>;

	// Nontemporal hints on the streams no other wave touches in this step:
	// the site's own slots, and the pulls and pushes with c_y = 0 (the +-1
	// shifted rows share cache lines with the neighbouring wave)
	static constexpr bool NT = (LBM_NT_GENERAL >> ST) & 1;
	static LBM_DEV real ldq(const elem* p, float sc, bool nt)
	{
		if (nt) {
			const elem v = ld_nt(p);
			return S::ld(&v, sc);
		}
		return S::ld(p, sc);
	}
	static LBM_DEV void stq(elem* p, real v, float inv, bool nt)
	{
		if (nt) {
			elem e;
			S::st(&e, v, inv);
			st_nt(p, e);
		}
		else
			S::st(p, v, inv);
	}

	static LBM_DEV int64_t sidx(const StepArgs& A, int64_t xs, int64_t ys, int64_t zs)
	{
		return (xs * A.Zs + zs) * A.Ys + ys;
	}
	template <typename E>
	static LBM_DEV E* pq(E* base, int q, const StepArgs& A, int64_t xs, int64_t ys, int64_t zs)
	{
		return slot<O32>(base, q, A, xs, ys, zs);
	}
	template <typename E>
	static LBM_DEV E* pq0(E* base, int q, const StepArgs& A, int64_t s0)
	{
		return slot0<O32>(base, q, A, s0);
	}

	static LBM_DEV void set_eq(real* f, real rho, real vx, real vy, real vz)
	{
#pragma unroll
		for (int q = 0; q < 27; q++)
			f[q] = eq_q(q, rho, vx, vy, vz);
	}
	static LBM_DEV real eq_q(int q, real rho, real vx, real vy, real vz)
	{
		// q is a compile-time constant after unrolling
		switch (q) {
#define EQC(Q) case Q: return P::template eq<EQK, Q>(rho, vx, vy, vz);
			EQC(0) EQC(1) EQC(2) EQC(3) EQC(4) EQC(5) EQC(6) EQC(7) EQC(8) EQC(9) EQC(10) EQC(11) EQC(12) EQC(13)
			EQC(14) EQC(15) EQC(16) EQC(17) EQC(18) EQC(19) EQC(20) EQC(21) EQC(22) EQC(23) EQC(24) EQC(25) EQC(26)
#undef EQC
		}
		return 0;
	}

	static LBM_DEV void inflow(const StepArgs& A, int64_t y, int64_t z, real& vx, real& vy, real& vz)
	{
		// NSE_Data_ConstInflow / NSE_Data_InflowProfile::inflow (lbm_data.h:85-132)
		const int64_t i = A.Y * z + y;
		vx = (real) inflow_component(A, 0, i);
		vy = (real) inflow_component(A, 1, i);
		vz = (real) inflow_component(A, 2, i);
	}

	static LBM_DEV void site(const StepArgs& A, int64_t x, int64_t y, int64_t z)
	{
		real vx, vy, vz;
		Nb NX, NY, NZ;
		site(A, x, y, z, vx, vy, vz, NX, NY, NZ);
	}

	// full-way bounce-back (bc.h:201-221)
	static LBM_DEV void bounce_back(real* f, real& rho, real& vx, real& vy, real& vz)
	{
		rho = 1;
		vx = vy = vz = 0;
#pragma unroll
		for (int q = 1; q < 27; q += 2) {
			const real t = f[q];
			f[q] = f[q + 1];
			f[q + 1] = t;
		}
	}

	// one site; returns the pre-collision velocity and the neighbour indices
	// (the coupled kernel drives the ADE step with them, kernels.h:152-176)
	static LBM_DEV void site(const StepArgs& A, int64_t x, int64_t y, int64_t z, real& ovx, real& ovy, real& ovz, Nb& NX,
							 Nb& NY, Nb& NZ)
	{
		const int64_t s0 = sidx(A, x + A.ox, y + A.oy, z + A.oz);
		const int16_t gi = A.map ? __builtin_nontemporal_load(A.map + s0) : (int16_t) GEO_PERIODIC;
		const bool per = gi == GEO_PERIODIC;
		NX = neighbours<AA>(x, A.X, A.ox, A.dist_x, per);
		NY = neighbours<AA>(y, A.Y, A.oy, A.dist_y, per);
		NZ = neighbours<AA>(z, A.Z, A.oz, A.dist_z, per);
		real Fx = (real) A.fx, Fy = (real) A.fy, Fz = (real) A.fz;
		if (A.force)
			add_site_force(A, s0, Fx, Fy, Fz);
		real f[27];
		real rho = 1, vx = 0, vy = 0, vz = 0;
		const elem* __restrict__ cur = (const elem*) A.df0;

		if (LEAN || gi != GEO_NOTHING) {
			// pull location (bc.h:72-74); OUTFLOW_RIGHT pulls as if located at x-1
			Nb PX = NX;
			if (! LEAN && gi == GEO_OUTFLOW_RIGHT)
				PX.p = PX.o = NX.m;
			if (! LEAN && gi == GEO_OUTFLOW_RIGHT_INTERP) {
				// streamingInterpRight (streaming_AB.h:446-478)
				const real cs = (real) 0.5773502691896257;
#pragma unroll
				for (int q = 0; q < 27; q++) {
					const int64_t yy = NY[-CY[q]], zz = NZ[-CZ[q]];
					if (CX[q] < 0)
						f[q] = cs * S::ld(cur + q * A.XYZ + sidx(A, NX.m, yy, zz), fx_sc(q, A.rs)) + ((real) 1 - cs) * S::ld(cur + q * A.XYZ + sidx(A, NX.o, yy, zz), fx_sc(q, A.rs));
					else
						f[q] = S::ld(cur + q * A.XYZ + sidx(A, CX[q] > 0 ? NX.m : NX.o, yy, zz), fx_sc(q, A.rs));
				}
			}
			else if (AA && A.even) {
#pragma unroll
				for (int q = 0; q < 27; q++)
					f[q] = ldq(pq(cur, q, A, PX.o, NY.o, NZ.o), fx_sc(q, A.rs), NT);
			}
			else if (AA) {
#pragma unroll
				for (int q = 0; q < 27; q++)
					f[q] = ldq(pq(cur, opp(q), A, PX[-CX[q]], NY[-CY[q]], NZ[-CZ[q]]), fx_sc(opp(q), A.rs),
							   NT && CY[q] == 0);
			}
			else {
#pragma unroll
				for (int q = 0; q < 27; q++)
					f[q] = ldq(pq(cur, q, A, PX[-CX[q]], NY[-CY[q]], NZ[-CZ[q]]), fx_sc(q, A.rs), NT && CY[q] == 0);
			}

			// the raw moments of the pulled populations, recorded after the step's own stores
			using TT = std::conditional_t<TAPE == 2, double, float>;
			TT tm[10];
			const bool rec = TAPE != 0 && (gi == GEO_FLUID || gi == GEO_PERIODIC || gi == GEO_OUTFLOW_RIGHT || gi == GEO_OUTFLOW_EQ);
			if constexpr (TAPE != 0) {
				real m[10];
				raw_moments10(f, m);
#pragma unroll
				for (int k = 0; k < 10; k++)
					tm[k] = (TT) m[k];
			}
			const real rest = S::dev ? (real) 1 : (real) 0;
			if constexpr (LEAN != 0) {
				if (LEAN == 2 && gi == GEO_WALL)
					bounce_back(f, rho, vx, vy, vz);
				else if (! FAST)
					P::density_velocity(f, Fx, Fy, Fz, rho, vx, vy, vz);
			}
			else switch (gi) {
				case GEO_INFLOW:
					inflow(A, y, z, vx, vy, vz);
					rho = 1;
					set_eq(f, rho, vx, vy, vz);
					break;
				case GEO_INFLOW_LEFT:
					if constexpr (! S::dev) {
						// moment-based inflow (bc.h:91-142)
						inflow(A, y, z, vx, vy, vz);
						rho = (real) 1 / (1 - vx)
							* ((f[zzz] + (((f[zpp] + f[zmm]) + (f[zpm] + f[zmp])) + ((f[zpz] + f[zmz]) + (f[zzp] + f[zzm]))))
							   + 2 * (f[mzz] + (((f[mpp] + f[mmm]) + (f[mpm] + f[mmp])) + ((f[mpz] + f[mmz]) + (f[mzp] + f[mzm])))));
						const real n1o3 = (real) (1.0 / 3.0), n1o9 = (real) (1.0 / 9.0);
						const real m100 = rho * vx, m010 = rho * vy, m001 = rho * vz, m011 = rho * (vy * vz);
						const real m020 = n1o3 * rho + rho * (vy * vy);
						const real m002 = n1o3 * rho + rho * (vz * vz);
						const real m021 = n1o3 * rho * vz + rho * ((vy * vy) * vz);
						const real m012 = n1o3 * rho * vy + rho * (vy * (vz * vz));
						const real m022 = n1o9 * rho + n1o3 * rho * (vy * vy + vz * vz) + rho * (vy * vy) * (vz * vz);
						f[pzz] = m100 + (m022 - (m020 + m002)) + f[mzz]
							   + (((f[zpp] + f[zmm]) + (f[zpm] + f[zmp])) + ((f[zzp] + f[zzm]) + (f[zpz] + f[zmz])))
							   + 2 * (((f[mpp] + f[mmm]) + (f[mpm] + f[mmp])) + ((f[mpz] + f[mmz]) + (f[mzp] + f[mzm])));
						f[ppz] = (real) 0.5 * ((m020 - m022) + (-m012 + m010)) - (f[mpz] + f[zpz]);
						f[pmz] = (real) 0.5 * ((m020 - m022) + (m012 - m010)) - (f[mmz] + f[zmz]);
						f[pzp] = (real) 0.5 * ((m002 - m022) + (-m021 + m001)) - (f[mzp] + f[zzp]);
						f[pzm] = (real) 0.5 * ((m002 - m022) + (m021 - m001)) - (f[mzm] + f[zzm]);
						f[ppp] = (real) 0.25 * ((m022 + m011) + (m021 + m012)) - (f[mpp] + f[zpp]);
						f[ppm] = (real) 0.25 * ((m022 - m011) + (-m021 + m012)) - (f[mpm] + f[zpm]);
						f[pmp] = (real) 0.25 * ((m022 - m011) + (m021 - m012)) - (f[mmp] + f[zmp]);
						f[pmm] = (real) 0.25 * ((m022 + m011) + (-m021 - m012)) - (f[mmm] + f[zmm]);
					}
					else
						P::density_velocity(f, Fx, Fy, Fz, rho, vx, vy, vz);
					break;
				case GEO_INFLOW_BB_LEFT:
					{
						// bc.h:143-165 (pull scheme): reads unstreamed own slots
						inflow(A, y, z, vx, vy, vz);
						const real k2 = (real) 2 * (real) 3;
						const int64_t so = sidx(A, NX.o, NY.o, NZ.o);
						f[pzz] = S::ld(cur + mzz * A.XYZ + so, fx_sc(mzz, A.rs)) - k2 * (real) (2.0 / 27.0) * rho * (-vx);
						f[ppz] = S::ld(cur + mmz * A.XYZ + so, fx_sc(mmz, A.rs)) - k2 * (real) (1.0 / 54.0) * rho * (-vx - vy);
						f[pmz] = S::ld(cur + mpz * A.XYZ + so, fx_sc(mpz, A.rs)) - k2 * (real) (1.0 / 54.0) * rho * (-vx + vy);
						f[pzp] = S::ld(cur + mzm * A.XYZ + so, fx_sc(mzm, A.rs)) - k2 * (real) (1.0 / 54.0) * rho * (-vx - vz);
						f[pzm] = S::ld(cur + mzp * A.XYZ + so, fx_sc(mzp, A.rs)) - k2 * (real) (1.0 / 54.0) * rho * (-vx + vz);
						f[ppp] = S::ld(cur + mmm * A.XYZ + so, fx_sc(mmm, A.rs)) - k2 * (real) (1.0 / 216.0) * rho * (-vx - vy - vz);
						f[ppm] = S::ld(cur + mmp * A.XYZ + so, fx_sc(mmp, A.rs)) - k2 * (real) (1.0 / 216.0) * rho * (-vx - vy + vz);
						f[pmp] = S::ld(cur + mpm * A.XYZ + so, fx_sc(mpm, A.rs)) - k2 * (real) (1.0 / 216.0) * rho * (-vx + vy - vz);
						f[pmm] = S::ld(cur + mpp * A.XYZ + so, fx_sc(mpp, A.rs)) - k2 * (real) (1.0 / 216.0) * rho * (-vx + vy + vz);
						P::density_velocity(f, Fx, Fy, Fz, rho, vx, vy, vz);
						break;
					}
				case GEO_INFLOW_EQ_LEFT:
					// bc.h:166-185 (deviation storage: the known-direction weights sum to 1)
					inflow(A, y, z, vx, vy, vz);
					rho = (real) 1 / (1 - vx)
						* ((rest + (f[zzz] + (((f[zpp] + f[zmm]) + (f[zpm] + f[zmp])) + ((f[zpz] + f[zmz]) + (f[zzp] + f[zzm])))))
						   + 2 * (f[mzz] + (((f[mpp] + f[mmm]) + (f[mpm] + f[mmp])) + ((f[mpz] + f[mmz]) + (f[mzp] + f[mzm])))));
					set_eq(f, rho, vx, vy, vz);
					break;
				case GEO_OUTFLOW_EQ:
					P::density_velocity(f, Fx, Fy, Fz, rho, vx, vy, vz);
					rho = 1;
					set_eq(f, rho, vx, vy, vz);
					break;
				case GEO_OUTFLOW_RIGHT:
					P::density_velocity(f, Fx, Fy, Fz, rho, vx, vy, vz);
					rho = 1;
					break;
				case GEO_OUTFLOW_RIGHT_INTERP:
					P::density_velocity(f, Fx, Fy, Fz, rho, vx, vy, vz);
#pragma unroll
					for (int q = 0; q < 27; q++)
						f[q] += eq_q(q, (real) 1, vx, vy, vz) - eq_q(q, rho, vx, vy, vz);
					rho = 1;
					break;
				case GEO_WALL:
					bounce_back(f, rho, vx, vy, vz);
					break;
				case GEO_SYM_TOP:
				case GEO_SYM_BOTTOM:
				case GEO_SYM_LEFT:
				case GEO_SYM_RIGHT:
				case GEO_SYM_BACK:
				case GEO_SYM_FRONT:
					{
						// bc.h:222-293
						const int axis = (gi == GEO_SYM_TOP || gi == GEO_SYM_BOTTOM) ? 2 : (gi == GEO_SYM_LEFT || gi == GEO_SYM_RIGHT) ? 0 : 1;
						const int sgn = (gi == GEO_SYM_BOTTOM || gi == GEO_SYM_LEFT || gi == GEO_SYM_BACK) ? 1 : -1;
#pragma unroll
						for (int q = 0; q < 27; q++) {
							const int c = axis == 0 ? CX[q] : axis == 1 ? CY[q] : CZ[q];
							if (c == sgn) {
								const int mq = axis == 0 ? qidx(1 - sgn, CY[q] + 1, CZ[q] + 1)
											 : axis == 1 ? qidx(CX[q] + 1, 1 - sgn, CZ[q] + 1)
														 : qidx(CX[q] + 1, CY[q] + 1, 1 - sgn);
								real src = 0;
#pragma unroll
								for (int r = 0; r < 27; r++)
									if (r == mq)
										src = f[r];
								f[q] = src;
							}
						}
						P::density_velocity(f, Fx, Fy, Fz, rho, vx, vy, vz);
						break;
					}
				default:
					// the fast FP32 collision sums its own moments at plain collided sites
					if (! (FAST && (gi == GEO_FLUID || gi == GEO_PERIODIC)))
						P::density_velocity(f, Fx, Fy, Fz, rho, vx, vy, vz);
					break;
			}
			// doCollision (bc.h:406-412)
			const bool do_coll = LEAN ? (LEAN == 1 || gi != GEO_WALL)
									  : (gi == GEO_FLUID || gi == GEO_PERIODIC || gi == GEO_OUTFLOW_RIGHT || gi == GEO_OUTFLOW_RIGHT_INTERP
										 || gi == GEO_INFLOW_LEFT || gi == GEO_INFLOW_BB_LEFT);
			if constexpr (FAST) {
				// FP32 forms, default cumulant: the raw-moment formulation (as the periodic kernel)
				// wherever rho and u are the site's own; the boundary codes that replace them keep
				// the reference chain
				if (LEAN ? do_coll : (gi == GEO_FLUID || gi == GEO_PERIODIC)) {
					using FC = FastCum<S::dev>;
					FC::collide(f, Fx, Fy, Fz, (real) A.omega1, true, rho, vx, vy, vz);
#pragma unroll
					for (int q = 0; q < 27; q++)
						f[q] *= FC::unscale(q);
				}
				else if (! LEAN && do_coll)
					collide<ST, COLL>(f, rho, vx, vy, vz, A, Fx, Fy, Fz);
			}
			else if (do_coll)
				collide<ST, COLL>(f, rho, vx, vy, vz, A, Fx, Fy, Fz);
			// postCollision streaming
			if (AA && A.even) {
				elem* __restrict__ out = (elem*) A.df0;
#pragma unroll
				for (int q = 0; q < 27; q++)
					stq(pq0(out, opp(q), A, s0), f[q], fx_inv(opp(q), A.invr), NT);
			}
			else if (AA) {
				elem* __restrict__ out = (elem*) A.df0;
#pragma unroll
				for (int q = 0; q < 27; q++)
					stq(pq(out, q, A, NX[CX[q]], NY[CY[q]], NZ[CZ[q]]), f[q], fx_inv(q, A.invr), NT && CY[q] == 0);
			}
			else {
				elem* __restrict__ out = (elem*) A.df1;
#pragma unroll
				for (int q = 0; q < 27; q++)
					stq(pq0(out, q, A, s0), f[q], fx_inv(q, A.invr), NT);
			}
			if constexpr (TAPE != 0) {
				if (rec) {
					TT* __restrict__ tp = (TT*) A.tape;
#pragma unroll
					for (int k = 0; k < 10; k++)
						__builtin_nontemporal_store(tm[k], tp + tape_index(k, s0));
				}
			}
		}
		if (A.out_macro)
			write_macro(A, s0, rho, vx, vy, vz);
		ovx = vx;
		ovy = vy;
		ovz = vz;
	}
};

// The FP64 default cumulant needs ~170 VGPRs here: asking for 3 waves per SIMD
// (<= 168) costs at most 2 spilled registers; the alternative operators keep
// their registers.
// LEAN 0: the box (or sub-box) of the launch, every code; 1, 2: a row list
// of that lean level; 3: a row list through the general paths
template <int ST, int COLL, int PATTERN, int TAPE = 0, bool O32 = false, int LEAN = 0>
__global__ void __launch_bounds__(256)
__attribute__((amdgpu_waves_per_eu(COLL == TNL_LBM_COLL_CUM ? 3 : 1, 8))) k_general(const StepArgs A)
{
	int64_t x, z, y;
	if (! block_coords<true, LEAN != 0>(A, x, z, y))
		return;
	General<ST, COLL, PATTERN, TAPE, O32, LEAN == 3 ? 0 : LEAN>::site(A, x, y, z);
}

// ===========================================================================
// Coupled NSE + ADE: cudaLBMKernel<NSE,ADE> (kernels.h:114-196)
// ===========================================================================
// The D3Q7 inputs of one site: its code, the 7 pulled populations, the
// diffusivity and the source.  Loaded before the D3Q27 update of the same
// site, so that their latency overlaps it (the compiler cannot move them above
// the D3Q27 stores by itself).
template <int ST>
struct AdeIn
{
	int16_t gi;
	typename Storage<ST>::real f[7];
	typename Storage<ST>::real D, S;
};

template <int ST, int PATTERN, bool O32 = false, int LEAN = 0>
LBM_DEV void ade_load(const StepArgs& A, const AdeArgs& B, int64_t s0, const Nb& NX, const Nb& NY, const Nb& NZ,
					  AdeIn<ST>& in)
{
	using K = Ade7<ST>;
	using real = typename K::real;
	using elem = typename K::elem;
	constexpr bool AA = PATTERN == TNL_LBM_PATTERN_AA;
	const int16_t gi = B.map ? B.map[s0] : (int16_t) A_PERIODIC;
	in.gi = gi;
	const real pr = (real) B.phi_ref, isc = (real) 1 / (real) B.scale;
	auto macro_at = [&](const void* p) -> real {
		return A.macro_f64 ? (real) ((const double*) p)[s0] : (real) ((const float*) p)[s0];
	};
	// D3Q7_MACRO_Default::copyQuantities (d3q7/macro.h:38-42)
	in.D = B.diff ? macro_at(B.diff) : (real) B.D;
	in.S = B.source ? macro_at(B.source) : (real) 0;
	if (LEAN || gi != A_NOTHING) {
		const elem* cur = (const elem*) B.df0;
		Nb PX = NX;
		if (! LEAN && gi == A_OUTFLOW_RIGHT)
			PX.p = PX.o = NX.m;
		// nontemporal hints as in the D3Q27 update: own slots and c_y = 0 pulls
		constexpr bool NT = (LBM_NT_GENERAL >> ST) & 1;
		auto ld = [&](const elem* p, int qs, bool nt) -> real {
			if (nt) {
				const elem v = ld_nt(p);
				return K::ld(&v, qs, pr, isc);
			}
			return K::ld(p, qs, pr, isc);
		};
#pragma unroll
		for (int q = 0; q < 7; q++) {
			if (AA && B.even)
				in.f[q] = ld(slot0<O32>(cur, q, A, s0), q, NT);
			else if (AA)
				in.f[q] = ld(slot<O32>(cur, opp(q), A, PX[-CX[q]], NY[-CY[q]], NZ[-CZ[q]]), opp(q), NT && CY[q] == 0);
			else
				in.f[q] = ld(slot<O32>(cur, q, A, PX[-CX[q]], NY[-CY[q]], NZ[-CZ[q]]), q, NT && CY[q] == 0);
		}
	}
}

// LEAN as in General: 1 = A_FLUID and A_PERIODIC sites only, 2 = those and A_WALL
template <int ST, int PATTERN, bool O32 = false, int LEAN = 0>
LBM_DEV void ade_site(const StepArgs& A, const AdeArgs& B, int64_t s0, typename Storage<ST>::real vx,
					  typename Storage<ST>::real vy, typename Storage<ST>::real vz, const Nb& NX, const Nb& NY, const Nb& NZ,
					  const AdeIn<ST>& in)
{
	using K = Ade7<ST>;
	using real = typename K::real;
	using elem = typename K::elem;
	constexpr bool AA = PATTERN == TNL_LBM_PATTERN_AA;
	const int16_t gi = in.gi;
	const real pr = (real) B.phi_ref, sc = (real) B.scale;
	const real D = in.D;
	real phi = 0;
	if (LEAN || gi != A_NOTHING) {
		real f[7];
#pragma unroll
		for (int q = 0; q < 7; q++)
			f[q] = in.f[q];
		if constexpr (LEAN != 0) {
			phi = K::density(f);
			if (LEAN == 2 && gi == A_WALL)
				K::bounce_back(f);
		}
		else switch (gi) {
			case A_INFLOW:
				phi = (real) B.inflow_phi;
				K::set_eq(f, phi, vx, vy, vz);
				break;
			case A_WALL:
				phi = K::density(f);
				K::bounce_back(f);
				break;
			case A_SYM_TOP: f[zzm] = f[zzp]; phi = K::density(f); break;
			case A_SYM_BOTTOM: f[zzp] = f[zzm]; phi = K::density(f); break;
			case A_SYM_LEFT: f[pzz] = f[mzz]; phi = K::density(f); break;
			case A_SYM_RIGHT: f[mzz] = f[pzz]; phi = K::density(f); break;
			case A_SYM_BACK: f[zpz] = f[zmz]; phi = K::density(f); break;
			case A_SYM_FRONT: f[zmz] = f[zpz]; phi = K::density(f); break;
			default: phi = K::density(f); break;
		}
		if (LEAN ? (LEAN == 1 || gi != A_WALL) : (gi == A_FLUID || gi == A_PERIODIC || gi == A_SOLID || gi == A_OUTFLOW_RIGHT)) {
			K::collide(B.collision, f, phi, vx, vy, vz, D);
			if (B.source) {
#pragma unroll
				for (int q = 0; q < 7; q++)
					f[q] += (real) w7(q) * in.S;
			}
		}
		elem* out = (elem*) (AA ? B.df0 : B.df1);
		constexpr bool NT = (LBM_NT_GENERAL >> ST) & 1;
		auto st = [&](elem* p, int qs, real v, bool nt) {
			if (nt) {
				elem e;
				K::st(&e, qs, v, pr, sc);
				st_nt(p, e);
			}
			else
				K::st(p, qs, v, pr, sc);
		};
#pragma unroll
		for (int q = 0; q < 7; q++) {
			if (AA && B.even)
				st(slot0<O32>(out, opp(q), A, s0), opp(q), f[q], NT);
			else if (AA)
				st(slot<O32>(out, q, A, NX[CX[q]], NY[CY[q]], NZ[CZ[q]]), q, f[q], NT && CY[q] == 0);
			else
				st(slot0<O32>(out, q, A, s0), q, f[q], NT);
		}
	}
	if (A.out_macro && B.macro) {
		if (A.macro_f64)
			((double*) B.macro)[s0] = phi;
		else
			((float*) B.macro)[s0] = (float) phi;
	}
}

// LEAN as for k_general
template <int ST, int COLL, int PATTERN, bool O32 = false, int LEAN = 0>
__global__ void __launch_bounds__(256) k_nse_ade(const StepArgs A, const AdeArgs B)
{
	constexpr int LV = LEAN == 3 ? 0 : LEAN;
	int64_t x, z, y;
	if (! block_coords<true, LEAN != 0>(A, x, z, y))
		return;
	constexpr bool AA = PATTERN == TNL_LBM_PATTERN_AA;
	const int64_t s0 = ((x + A.ox) * A.Zs + (z + A.oz)) * A.Ys + (y + A.oy);
	// neighbours as General::site forms them (periodic wrap for PERIODIC-coded sites)
	const bool per = (A.map ? A.map[s0] : (int16_t) GEO_PERIODIC) == GEO_PERIODIC;
	const Nb NX = neighbours<AA>(x, A.X, A.ox, A.dist_x, per);
	const Nb NY = neighbours<AA>(y, A.Y, A.oy, A.dist_y, per);
	const Nb NZ = neighbours<AA>(z, A.Z, A.oz, A.dist_z, per);
	AdeIn<ST> in;
	ade_load<ST, PATTERN, O32, LV>(A, B, s0, NX, NY, NZ, in);
	typename Storage<ST>::real vx, vy, vz;
	Nb GX, GY, GZ;
	General<ST, COLL, PATTERN, 0, O32, LV>::site(A, x, y, z, vx, vy, vz, GX, GY, GZ);
	ade_site<ST, PATTERN, O32, LV>(A, B, s0, vx, vy, vz, NX, NY, NZ, in);
}

template <int ST>
__global__ void k_ade_set_equilibrium(void* df0, void* df1, int ncopies, int64_t XYZ, int64_t n, const void* phi,
									  int phi_f64, double phi0, double pr, double sc)
{
	using K = Ade7<ST>;
	const int64_t s = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
	if (s >= n)
		return;
	const double p = phi ? (phi_f64 ? ((const double*) phi)[s] : (double) ((const float*) phi)[s]) : phi0;
	typename K::real f[7];
	K::set_eq(f, (typename K::real) p, 0, 0, 0);
	for (int k = 0; k < ncopies; k++) {
		auto* out = (typename K::elem*) (k == 0 ? df0 : df1);
#pragma unroll
		for (int q = 0; q < 7; q++)
			K::st(out + q * XYZ + s, q, f[q], (typename K::real) pr, (typename K::real) sc);
	}
}

template <int ST>
__global__ void k_ade_export(const void* df, double* out, int64_t XYZ, int64_t n, double pr, double sc)
{
	using K = Ade7<ST>;
	const int64_t s = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
	if (s >= n)
		return;
#pragma unroll
	for (int q = 0; q < 7; q++)
		out[q * XYZ + s] = (double) K::ld((const typename K::elem*) df + q * XYZ + s, q, (typename K::real) pr,
										  (typename K::real) 1 / (typename K::real) sc);
}

// ===========================================================================
// Row classification for the row-list launches (tnl_lbm_classify_rows)
// ===========================================================================
// One wave per interior (x, z) row, row i = x * Z + z: class 1 when every
// code of the row is FLUID or PERIODIC (D3Q27 and, with amap, D3Q7), 2 when
// WALL sites occur as well, 0 otherwise.  A null map is all PERIODIC.
static __global__ void __launch_bounds__(256) k_classify_rows(const int16_t* map, const int16_t* amap, int64_t X, int64_t Y,
													   int64_t Z, int64_t ox, int64_t oy, int64_t oz, int64_t Ys, int64_t Zs,
													   uint8_t* cls)
{
	const int64_t i = ((int64_t) blockIdx.x * blockDim.x + threadIdx.x) >> 6;  // uniform per wave
	if (i >= X * Z)
		return;
	const int64_t x = i / Z, z = i % Z;
	const int64_t base = ((x + ox) * Zs + (z + oz)) * Ys + oy;
	bool other = false, wall = false;
	for (int64_t y = threadIdx.x & 63; y < Y; y += 64) {
		if (map) {
			const int16_t c = map[base + y];
			wall |= c == GEO_WALL;
			other |= ! (c == GEO_FLUID || c == GEO_PERIODIC || c == GEO_WALL);
		}
		if (amap) {
			const int16_t c = amap[base + y];
			wall |= c == A_WALL;
			other |= ! (c == A_FLUID || c == A_PERIODIC || c == A_WALL);
		}
	}
	const uint64_t o = __ballot(other), w = __ballot(wall);
	if ((threadIdx.x & 63) == 0)
		cls[i] = o ? 0 : (w ? 2 : 1);
}

// The three row lists from the classes, in one workgroup so that each list
// keeps increasing row order: counts first, then every chunk of 1024 rows
// places its rows of each class after those of the earlier chunks.
static __global__ void __launch_bounds__(1024) k_row_lists(const uint8_t* cls, uint32_t n, uint32_t Z, uint32_t* rows,
													uint32_t* counts)
{
	__shared__ uint32_t part[16][3];
	__shared__ uint32_t off[3];
	const uint32_t t = threadIdx.x, lane = t & 63, wave = t >> 6;
	uint32_t c3[3] = {0, 0, 0};
	for (uint32_t i = t; i < n; i += 1024) {
		const uint8_t c = cls[i];
		c3[0] += c == 0;
		c3[1] += c == 1;
		c3[2] += c == 2;
	}
	for (int k = 0; k < 3; k++) {
		uint32_t v = c3[k];
		for (int d = 32; d >= 1; d >>= 1)
			v += __shfl_xor(v, d);
		if (lane == 0)
			part[wave][k] = v;
	}
	__syncthreads();
	if (t == 0) {
		uint32_t tot[3] = {0, 0, 0};
		for (int w = 0; w < 16; w++)
			for (int k = 0; k < 3; k++)
				tot[k] += part[w][k];
		for (int k = 0; k < 3; k++)
			counts[k] = tot[k];
		off[0] = 0;
		off[1] = tot[0];
		off[2] = tot[0] + tot[1];
	}
	__syncthreads();
	const uint64_t below = (lane == 0) ? 0 : (~0ull >> (64 - lane));
	for (uint32_t c0 = 0; c0 < n; c0 += 1024) {
		const uint32_t i = c0 + t;
		const int c = i < n ? (int) cls[i] : -1;
		uint64_t m[3];
		for (int k = 0; k < 3; k++) {
			m[k] = __ballot(c == k);
			if (lane == 0)
				part[wave][k] = (uint32_t) __popcll(m[k]);
		}
		__syncthreads();
		if (c >= 0) {
			uint32_t pos = off[c];
			for (uint32_t w = 0; w < wave; w++)
				pos += part[w][c];
			const uint64_t mc = c == 0 ? m[0] : (c == 1 ? m[1] : m[2]);
			pos += (uint32_t) __popcll(mc & below);
			rows[pos] = ((i / Z) << 16) | (i % Z);
		}
		__syncthreads();
		if (t < 3) {
			uint32_t add = 0;
			for (int w = 0; w < 16; w++)
				add += part[w][t];
			off[t] += add;
		}
		__syncthreads();
	}
}

// ===========================================================================
// Initialisation, transfer and reduction kernels
// ===========================================================================
struct FieldArgs
{
	void* df[2];
	int ncopies;
	int64_t XYZ;
	int64_t nsites;  // stored sites incl. overlaps
	const double* rho;
	const double* vx;
	const double* vy;
	const double* vz;
	double c[4];
	float rs[4], invr[4];
};

template <int ST, int EQK>
__global__ void k_set_equilibrium(const FieldArgs A)
{
	using S = Storage<ST>;
	using real = typename S::real;
	const int64_t s = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
	if (s >= A.nsites)
		return;
	const real rho = (real) (A.rho ? A.rho[s] : A.c[0]);
	const real vx = (real) (A.vx ? A.vx[s] : A.c[1]);
	const real vy = (real) (A.vy ? A.vy[s] : A.c[2]);
	const real vz = (real) (A.vz ? A.vz[s] : A.c[3]);
	real f[27];
	// one operator per equilibrium kind
	constexpr int coll = EQK == 0 ? TNL_LBM_COLL_CUM : EQK == 1 ? TNL_LBM_COLL_MRT_LES : TNL_LBM_COLL_KBC_N2;
	General<ST, coll, TNL_LBM_PATTERN_AA>::set_eq(f, rho, vx, vy, vz);
	for (int k = 0; k < A.ncopies; k++) {
		auto* out = (typename S::elem*) A.df[k];
#pragma unroll
		for (int q = 0; q < 27; q++)
			S::st(out + q * A.XYZ + s, f[q], fx_inv(q, A.invr));
	}
}

struct MacroArgs
{
	float rs[4], invr[4];
	const void* df;
	void* macro;
	int macro_f64;
	int64_t XYZ;
	int64_t X, Y, Z, ox, oy, oz, Ys, Zs;
	int32_t dist_x, dist_y, dist_z;  // gradient reduction: neighbours from the overlaps, not wrapped
};

// LBM_BLOCK::computeInitialMacro (lbm_block.hpp:281-305); forces zeroed
template <int ST>
__global__ void k_compute_macro(const MacroArgs A)
{
	using S = Storage<ST>;
	using real = typename S::real;
	const int64_t n = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
	if (n >= A.X * A.Y * A.Z)
		return;
	const int64_t y = n % A.Y, t = n / A.Y, z = t % A.Z, x = t / A.Z;
	const int64_t s = ((x + A.ox) * A.Zs + (z + A.oz)) * A.Ys + (y + A.oy);
	real f[27];
#pragma unroll
	for (int q = 0; q < 27; q++)
		f[q] = S::ld((const typename S::elem*) A.df + q * A.XYZ + s, fx_sc(q, A.rs));
	real rho, vx, vy, vz;
	Phys<ST>::density_velocity(f, 0, 0, 0, rho, vx, vy, vz);
	if (A.macro_f64) {
		double* m = (double*) A.macro;
		m[s] = rho; m[A.XYZ + s] = vx; m[2 * A.XYZ + s] = vy; m[3 * A.XYZ + s] = vz;
	}
	else {
		float* m = (float*) A.macro;
		m[s] = rho; m[A.XYZ + s] = vx; m[2 * A.XYZ + s] = vy; m[3 * A.XYZ + s] = vz;
	}
}

template <int ST, int Q, bool EXPORT>
__global__ void k_transfer(void* df, double* buf, int64_t XYZ, int64_t nsites, const FixedRange R)
{
	using S = Storage<ST>;
	const int64_t s = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
	if (s >= nsites)
		return;
	auto* p = (typename S::elem*) df;
#pragma unroll
	for (int q = 0; q < Q; q++) {
		const double w = S::dev ? (Q == 27 ? w27(q) : (q == 0 ? 0.25 : 0.125)) : 0.0;
		if constexpr (EXPORT)
			buf[q * XYZ + s] = (double) S::ld(p + q * XYZ + s, fx_sc(q, R.rs)) + w;
		else
			S::st(p + q * XYZ + s, (typename S::real) (buf[q * XYZ + s] - w), fx_inv(q, R.invr));
	}
}

// FP64 reductions over the macro field: deterministic two-pass block sums
constexpr int RED_N = 6;
constexpr int RED_BLOCKS = 1024;

__device__ void block_sum(double* v, double (*sh)[RED_N])
{
	const int t = threadIdx.x;
#pragma unroll
	for (int i = 0; i < RED_N; i++)
		sh[t][i] = v[i];
	__syncthreads();
	for (int w = blockDim.x / 2; w > 0; w >>= 1) {
		if (t < w)
#pragma unroll
			for (int i = 0; i < RED_N; i++)
				sh[t][i] += sh[t + w][i];
		__syncthreads();
	}
}

static __global__ void __launch_bounds__(256) k_reduce_macro(const MacroArgs A, double* partial)
{
	__shared__ double sh[256][RED_N];
	double v[RED_N] = {0, 0, 0, 0, 0, 0};
	const int64_t N = A.X * A.Y * A.Z;
	for (int64_t n = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; n < N; n += (int64_t) gridDim.x * blockDim.x) {
		const int64_t y = n % A.Y, t = n / A.Y, z = t % A.Z, x = t / A.Z;
		const int64_t s = ((x + A.ox) * A.Zs + (z + A.oz)) * A.Ys + (y + A.oy);
		double rho, vx, vy, vz;
		if (A.macro_f64) {
			const double* m = (const double*) A.macro;
			rho = m[s]; vx = m[A.XYZ + s]; vy = m[2 * A.XYZ + s]; vz = m[3 * A.XYZ + s];
		}
		else {
			const float* m = (const float*) A.macro;
			rho = m[s]; vx = m[A.XYZ + s]; vy = m[2 * A.XYZ + s]; vz = m[3 * A.XYZ + s];
		}
		const double u2 = vx * vx + vy * vy + vz * vz;
		v[0] += rho;
		v[1] += rho * u2;
		v[2] += (rho != rho) ? 1.0 : 0.0;
		v[3] += (rho - 1.0) * (rho - 1.0);
		v[4] += u2;
		v[5] += 1.0;
	}
	block_sum(v, sh);
	if (threadIdx.x == 0)
#pragma unroll
		for (int i = 0; i < RED_N; i++)
			partial[blockIdx.x * RED_N + i] = sh[0][i];
}

// Velocity-gradient reductions of sim_4.cu:282-300: with the central
// differences G_ab = (v_a(r + e_b) - v_a(r - e_b)) / 2 of computeGradVel
// (sim_4.cu:20-60; periodic wrap on an axis that is not distributed, the
// overlap layer on one that is), per block:
//   [0] sum rho |G|^2       (the enstrophy integrand of sim_4.cu:284-298)
//   [1] sum rho |curl v|^2  (computeVorticity, sim_4.cu:62-82)
// entries 2..5 of each partial row are zero.
static __global__ void __launch_bounds__(256) k_reduce_gradients(const MacroArgs A, double* partial)
{
	__shared__ double sh[256][RED_N];
	double v[RED_N] = {0, 0, 0, 0, 0, 0};
	const int64_t N = A.X * A.Y * A.Z;
	auto at = [&](int id, int64_t x, int64_t y, int64_t z) -> double {
		const int64_t s = (int64_t) id * A.XYZ + ((x + A.ox) * A.Zs + (z + A.oz)) * A.Ys + (y + A.oy);
		return A.macro_f64 ? ((const double*) A.macro)[s] : (double) ((const float*) A.macro)[s];
	};
	for (int64_t n = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; n < N; n += (int64_t) gridDim.x * blockDim.x) {
		const int64_t y = n % A.Y, t = n / A.Y, z = t % A.Z, x = t / A.Z;
		const int64_t xp = (! A.dist_x && x == A.X - 1) ? 0 : x + 1, xm = (! A.dist_x && x == 0) ? A.X - 1 : x - 1;
		const int64_t yp = (! A.dist_y && y == A.Y - 1) ? 0 : y + 1, ym = (! A.dist_y && y == 0) ? A.Y - 1 : y - 1;
		const int64_t zp = (! A.dist_z && z == A.Z - 1) ? 0 : z + 1, zm = (! A.dist_z && z == 0) ? A.Z - 1 : z - 1;
		double G[3][3];  // G[a][b] = d v_a / d x_b
#pragma unroll
		for (int a = 0; a < 3; a++) {
			G[a][0] = 0.5 * (at(1 + a, xp, y, z) - at(1 + a, xm, y, z));
			G[a][1] = 0.5 * (at(1 + a, x, yp, z) - at(1 + a, x, ym, z));
			G[a][2] = 0.5 * (at(1 + a, x, y, zp) - at(1 + a, x, y, zm));
		}
		const double rho = at(0, x, y, z);
		double g2 = 0;
#pragma unroll
		for (int a = 0; a < 3; a++)
#pragma unroll
			for (int b = 0; b < 3; b++)
				g2 += G[a][b] * G[a][b];
		const double wx = G[2][1] - G[1][2], wy = G[0][2] - G[2][0], wz = G[1][0] - G[0][1];
		v[0] += rho * g2;
		v[1] += rho * (wx * wx + wy * wy + wz * wz);
	}
	block_sum(v, sh);
	if (threadIdx.x == 0)
#pragma unroll
		for (int i = 0; i < RED_N; i++)
			partial[blockIdx.x * RED_N + i] = sh[0][i];
}

static __global__ void __launch_bounds__(256) k_reduce_final(const double* partial, int nparts, double* out)
{
	__shared__ double sh[256][RED_N];
	double v[RED_N] = {0, 0, 0, 0, 0, 0};
	for (int b = threadIdx.x; b < nparts; b += blockDim.x)
#pragma unroll
		for (int i = 0; i < RED_N; i++)
			v[i] += partial[b * RED_N + i];
	block_sum(v, sh);
	if (threadIdx.x == 0)
#pragma unroll
		for (int i = 0; i < RED_N; i++)
			out[i] = sh[0][i];
}

// ===========================================================================
// Host side: validation and dispatch
// ===========================================================================
// Kernels that compute every lane's coordinates (k_general, k_nse_ade, the
// adjoint) can take the y-z plane of one x as one linear range, so that a row
// length that is not a multiple of the block (Y = 96, say) still fills every
// wave.  The map-free periodic kernel keeps whole rows per block.
inline void linearize(StepArgs& A, unsigned& threads, unsigned& grid)
{
	if (A.nx == 0 || A.ny % (int32_t) threads == 0)
		return;
	threads = 256;
	const uint64_t plane = (uint64_t) A.ny * (uint64_t) A.nz;
	A.nyb = (uint32_t) ((plane + threads - 1) / threads);
	A.nblocks = A.nyb * (uint32_t) A.nx;  // no larger than the row mapping's count
	const uint32_t padded = (A.nblocks + NXCD - 1) / NXCD * NXCD;
	A.per_xcd = padded / NXCD;
	A.lin = 1;
	grid = padded;
}

// Launch shape over one row list of a full-box launch: whole rows per block
// where the row length is a multiple of the block, else the listed rows'
// sites one after another, as linearize does for the box.
inline void list_shape(StepArgs& A, const uint32_t* rows, uint32_t nrows, unsigned& threads, unsigned& grid)
{
	A.rows = rows;
	A.nrows = nrows;
	threads = A.ny > 128 ? 256 : (A.ny > 64 ? 128 : 64);
	if (A.ny % (int32_t) threads == 0) {
		A.lin = 0;
		A.nyb = (uint32_t) A.ny / threads;
		A.nblocks = A.nyb * nrows;
	}
	else {
		threads = 256;
		A.lin = 1;
		A.nyb = 1;
		A.nblocks = (uint32_t) (((uint64_t) nrows * (uint64_t) A.ny + threads - 1) / threads);
	}
	const uint32_t padded = (A.nblocks + NXCD - 1) / NXCD * NXCD;
	A.per_xcd = padded / NXCD;
	grid = padded;
}

// The row lists of a full-box step: one launch per non-empty list, the lean
// kernels (LEAN 1, 2) for classes 1 and 2 and the general paths over a list
// (LEAN 3) for class 0.  go(B, threads, grid, lean) makes one launch.
template <typename Go>
int for_row_lists(const StepArgs& A, Go&& go)
{
	int64_t off = 0;
	for (int c = 0; c < 3; c++) {
		const int64_t n = A.list_count[c];
		if (n > 0) {
			StepArgs B = A;
			unsigned threads = 0, grid = 0;
			list_shape(B, A.lists + off, (uint32_t) n, threads, grid);
			if (int rc = go(B, threads, grid, c == 0 ? 3 : c))
				return rc;
		}
		off += n;
	}
	return 0;
}

// Launch shape of k_aa_even_pair (one lane per two sites of a row); false
// where the pairs would not be 4-byte aligned.
inline bool pair_shape(const StepArgs& A, StepArgs& B, unsigned& threads, unsigned& grid)
{
	if (((A.ny | A.offy | A.oy | A.Y | A.XYZ) & 1) || ((uintptr_t) A.df0 & 3))
		return false;
	B = A;
	const uint32_t np = (uint32_t) A.ny / 2;
	threads = np > 128 ? 256 : (np > 64 ? 128 : 64);
	B.nyb = (np + threads - 1) / threads;
	B.nblocks = B.nyb * (uint32_t) A.nz * (uint32_t) A.nx;  // no more blocks than one lane per site needs
	const uint32_t padded = (B.nblocks + NXCD - 1) / NXCD * NXCD;
	B.per_xcd = padded / NXCD;
	grid = padded;
	return true;
}

// The map-free periodic A-A kernels: the paired even step for the 2-byte
// forms where the row allows it, else one site per lane.
template <int ST, int COLL>
int launch_periodic(const StepArgs& A, unsigned threads, unsigned grid, hipStream_t st)
{
	if constexpr (sizeof(typename Storage<ST>::elem) == 2 && ((LBM_PAIR_EVEN >> ST) & 1)) {
		StepArgs B;
		unsigned t2 = 0, g2 = 0;
		if (A.even && pair_shape(A, B, t2, g2)) {
			hipLaunchKernelGGL((k_aa_even_pair<ST, COLL>), dim3(g2), dim3(t2), 0, st, B);
			return hip_check(hipGetLastError(), "step launch");
		}
	}
	if (ST != TNL_LBM_STORE_F64 && LBM_TPB_FP32 > 256 && A.ny % LBM_TPB_FP32 == 0) {
		StepArgs B = A;
		B.nyb = (uint32_t) A.ny / LBM_TPB_FP32;
		B.nblocks = B.nyb * (uint32_t) A.nz * (uint32_t) A.nx;  // fewer blocks than at 256
		const uint32_t padded = (B.nblocks + NXCD - 1) / NXCD * NXCD;
		B.per_xcd = padded / NXCD;
		if (A.even)
			hipLaunchKernelGGL((k_aa_periodic<ST, COLL, true>), dim3(padded), dim3(LBM_TPB_FP32), 0, st, B);
		else
			hipLaunchKernelGGL((k_aa_periodic<ST, COLL, false>), dim3(padded), dim3(LBM_TPB_FP32), 0, st, B);
	}
	else if (A.even)
		hipLaunchKernelGGL((k_aa_periodic<ST, COLL, true>), dim3(grid), dim3(threads), 0, st, A);
	else
		hipLaunchKernelGGL((k_aa_periodic<ST, COLL, false>), dim3(grid), dim3(threads), 0, st, A);
	return hip_check(hipGetLastError(), "step launch");
}

// One launch of the map kernel at lean level LEAN (k_general)
template <int ST, int COLL, int LEAN>
void launch_general(int pattern, const StepArgs& B, unsigned threads, unsigned grid, hipStream_t st, bool o32)
{
	if (o32) {
		if constexpr (sizeof(typename Storage<ST>::elem) == 2) {
			if (pattern == TNL_LBM_PATTERN_AA)
				hipLaunchKernelGGL((k_general<ST, COLL, TNL_LBM_PATTERN_AA, 0, true, LEAN>), dim3(grid), dim3(threads), 0, st, B);
			else
				hipLaunchKernelGGL((k_general<ST, COLL, TNL_LBM_PATTERN_AB, 0, true, LEAN>), dim3(grid), dim3(threads), 0, st, B);
		}
	}
	else if (pattern == TNL_LBM_PATTERN_AA)
		hipLaunchKernelGGL((k_general<ST, COLL, TNL_LBM_PATTERN_AA, 0, false, LEAN>), dim3(grid), dim3(threads), 0, st, B);
	else
		hipLaunchKernelGGL((k_general<ST, COLL, TNL_LBM_PATTERN_AB, 0, false, LEAN>), dim3(grid), dim3(threads), 0, st, B);
}

template <int ST, int COLL>
int launch_step(const tnl_lbm_config* cfg, const StepArgs& A, unsigned threads, unsigned grid, hipStream_t st)
{
	// the map-free periodic kernels take the homogeneous force only
	if (! A.map && ! A.force && cfg->pattern == TNL_LBM_PATTERN_AA)
		return launch_periodic<ST, COLL>(A, threads, grid, st);
	// 2-byte forms: 32-bit slot offsets where a population array is under
	// 4 GiB (F16 384^3 channel: A-A 1821 -> 1430 us, 139 -> 106 VGPRs;
	// A-B -2%; neutral for the 4- and 8-byte forms, which keep 64-bit)
	bool o32 = false;
	if constexpr (sizeof(typename Storage<ST>::elem) == 2)
		o32 = narrow_slots(A.XYZ, sizeof(typename Storage<ST>::elem));
	// row lists for the deviation forms, whose map kernel is bound by its
	// registers and instruction issue (384^3 with z walls, lean against
	// general: I16 -22%, F16 A-B -8%, BF16 -6%, F32_DEV -2%); the F64 and F32
	// map kernels gain nothing, and a class-0 list costs them an extra launch
	// (C4 duct +4 to +6%)
	if constexpr (COLL == TNL_LBM_COLL_CUM && Storage<ST>::dev) {
		if (A.lists)
			return for_row_lists(A, [&](const StepArgs& B, unsigned t, unsigned g, int lean) {
				if (lean == 1)
					launch_general<ST, COLL, 1>(cfg->pattern, B, t, g, st, o32);
				else if (lean == 2)
					launch_general<ST, COLL, 2>(cfg->pattern, B, t, g, st, o32);
				else
					launch_general<ST, COLL, 3>(cfg->pattern, B, t, g, st, o32);
				return hip_check(hipGetLastError(), "step launch");
			});
	}
	StepArgs B = A;
	linearize(B, threads, grid);
	launch_general<ST, COLL, 0>(cfg->pattern, B, threads, grid, st, o32);
	return hip_check(hipGetLastError(), "step launch");
}

// launches of the alternative operators (MRT_LES, KBC), built in lbm_alt.hip
int launch_alt(const tnl_lbm_config* cfg, const StepArgs& A, unsigned threads, unsigned grid, hipStream_t st);

}  // namespace lbm
