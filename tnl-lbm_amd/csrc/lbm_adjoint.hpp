// Adjoint (transposed Jacobian) of one A-B lattice step with the default
// cumulant collision, for the gradient of an objective with respect to the
// state and to the inflow velocity.
//
// The reference's adjoint (col_srt_adjoint.h, streaming_AB.h:246-281,
// bc.h:295-398) is for the SRT operator; this one is derived for the default
// cumulant (col_cum.h:13-511).  Key structure: the post-collision state of
// the default cumulant depends on the pre-collision populations only through
// ten raw moments (m0, j, Pi): rho and u, the second-order central moments
// (the only relaxed cumulants), and the 4th/6th-order products built from
// them.  So J^T lam = M^T (dG/dM)^T lam with M the constant 10 x 27 moment
// matrix, and (dG/dM)^T lam is one reverse sweep through the backward
// central-moment chain (27 columns of 3) and a few dozen scalar products --
// about the cost of the forward collision, so the adjoint step stays bound
// by memory traffic.
#pragma once

#include "lbm_step.hpp"

namespace lbm {

template <typename real>
struct AdjCum
{
	static constexpr int NMOM = 10;
	static LBM_DEV void raw_moments(const real* f, real* m) { raw_moments10(f, m); }

	// forward backward-pass of one column, absolute form:
	//   fz = k0 (1 - v^2) - 2 v k1 - k2,  f-+ = (k0 (v^2 -+ v) + k1 (2 v -+ 1) + k2) / 2
	static LBM_DEV void bwd(real k0, real k1, real k2, real v, real& fm, real& fz, real& fp)
	{
		const real v2 = v * v;
		fz = k0 * ((real) 1 - v2) - (real) 2 * v * k1 - k2;
		fm = (k0 * (v2 - v) + k1 * ((real) 2 * v - (real) 1) + k2) * (real) 0.5;
		fp = (k0 * (v2 + v) + k1 * ((real) 2 * v + (real) 1) + k2) * (real) 0.5;
	}
	// its adjoint: (lm, lz, lp) -> (l0, l1, l2), and d/dv accumulated into lv
	static LBM_DEV void bwd_adj(real k0, real k1, real v, real lm, real lz, real lp, real& l0, real& l1, real& l2, real& lv)
	{
		const real S = (real) 0.5 * (lm + lp);
		const real d = (real) 0.5 * (lp - lm);
		l2 = S - lz;
		l1 = (real) 2 * v * l2 + d;
		l0 = v * (v * l2 + d) + lz;
		lv += k0 * l1 + (real) 2 * k1 * l2;
	}

	// mu = (d f_out / d f_in)^T lam for the collision at the pre-collision
	// populations f.  rho_fixed: the caller has replaced rho by 1 before the
	// collision (OUTFLOW_RIGHT, bc.h:191-194): the 4th/6th-order products then
	// divide by 1 while u and the trace keep the actual mass.
	template <typename L, typename Mu>
	static LBM_DEV void vjp(const real* f, const L* lam, real Fx, real Fy, real Fz, real omega1, bool rho_fixed, Mu* mu)
	{
		real m[NMOM];
		raw_moments(f, m);
		vjp_m(m, lam, Fx, Fy, Fz, omega1, rho_fixed, mu);
	}

	// the same from the raw moments of the pre-collision populations, which
	// are all the Jacobian of the default cumulant depends on.  lam and mu may
	// be held in a narrower type than the arithmetic (float fields, double
	// sweep): each value is widened where it is read and rounded once where
	// it is written.
	template <typename L, typename Mu>
	static LBM_DEV void vjp_m(const real* m, const L* lam, real Fx, real Fy, real Fz, real omega1, bool rho_fixed, Mu* mu)
	{
		// ---- forward recomputation ---------------------------------------
		const real m0 = m[0], jx = m[1], jy = m[2], jz = m[3], Pxx = m[4], Pyy = m[5], Pzz = m[6], Pxy = m[7], Pxz = m[8],
				   Pyz = m[9];
		const real hx = (real) 0.5 * Fx, hy = (real) 0.5 * Fy, hz = (real) 0.5 * Fz;
		const real rv = (real) 1 / m0;
		const real vx = (jx + hx) * rv, vy = (jy + hy) * rv, vz = (jz + hz) * rv;
		const real rk = rho_fixed ? (real) 1 : rv;
		const real k200 = Pxx - vx * ((real) 2 * jx - m0 * vx);
		const real k020 = Pyy - vy * ((real) 2 * jy - m0 * vy);
		const real k002 = Pzz - vz * ((real) 2 * jz - m0 * vz);
		const real k110 = Pxy - vx * jy - vy * (jx - m0 * vx);
		const real k101 = Pxz - vx * jz - vz * (jx - m0 * vx);
		const real k011 = Pyz - vy * jz - vz * (jy - m0 * vy);
		const real om = (real) 1 - omega1;
		const real third = (real) (1.0 / 3.0);
		const real E33 = om * (k200 - k020), E34 = om * (k200 - k002), E35 = m0;
		const real a = third * (E33 + E34 + E35);
		const real b = third * (-(real) 2 * E33 + E34 + E35);
		const real c = third * (E33 - (real) 2 * E34 + E35);
		const real s = om * k110, p = om * k101, qq = om * k011;
		const real X211 = a * qq + (real) 2 * p * s, X121 = b * p + (real) 2 * s * qq, X112 = c * s + (real) 2 * qq * p;
		const real X220 = b * a + (real) 2 * s * s, X022 = c * b + (real) 2 * qq * qq, X202 = a * c + (real) 2 * p * p;
		const real ks211 = X211 * rk, ks121 = X121 * rk, ks112 = X112 * rk;
		const real ks220 = X220 * rk, ks022 = X022 * rk, ks202 = X202 * rk;
		const real Y1 = a * ks022 + b * ks202 + c * ks220 + (real) 4 * (qq * ks211 + p * ks121 + s * ks112);
		const real Y2 = (real) 16 * s * p * qq + (real) 4 * (p * p * b + qq * qq * a + s * s * c) + (real) 2 * a * b * c;
		const real ks222 = Y1 * rk - Y2 * rk * rk;
		// x-pass column inputs K[j][k][m] (moment m along x; j, k along y, z)
		real K[3][3][3];
		K[0][0][0] = m0;    K[0][0][1] = hx;    K[0][0][2] = a;
		K[1][0][0] = hy;    K[1][0][1] = s;     K[1][0][2] = 0;
		K[2][0][0] = b;     K[2][0][1] = 0;     K[2][0][2] = ks220;
		K[0][1][0] = hz;    K[0][1][1] = p;     K[0][1][2] = 0;
		K[1][1][0] = qq;    K[1][1][1] = 0;     K[1][1][2] = ks211;
		K[2][1][0] = 0;     K[2][1][1] = ks121; K[2][1][2] = 0;
		K[0][2][0] = c;     K[0][2][1] = 0;     K[0][2][2] = ks202;
		K[1][2][0] = 0;     K[1][2][1] = ks112; K[1][2][2] = 0;
		K[2][2][0] = ks022; K[2][2][1] = 0;     K[2][2][2] = ks222;
		real U[3][3][3], T[3][3][3];  // U[i][j][k]: x position i; T[i][j][k]: x, y positions i, j
#pragma unroll
		for (int j = 0; j < 3; j++)
#pragma unroll
			for (int k = 0; k < 3; k++)
				bwd(K[j][k][0], K[j][k][1], K[j][k][2], vx, U[0][j][k], U[1][j][k], U[2][j][k]);
#pragma unroll
		for (int i = 0; i < 3; i++)
#pragma unroll
			for (int k = 0; k < 3; k++)
				bwd(U[i][0][k], U[i][1][k], U[i][2][k], vy, T[i][0][k], T[i][1][k], T[i][2][k]);

		// ---- reverse sweep ---------------------------------------------------
		real lvx = 0, lvy = 0, lvz = 0;
		real lT[3][3][3], lU[3][3][3], lK[3][3][3];
#pragma unroll
		for (int i = 0; i < 3; i++)
#pragma unroll
			for (int j = 0; j < 3; j++)
				bwd_adj(T[i][j][0], T[i][j][1], vz, lam[qidx(i, j, 0)], lam[qidx(i, j, 1)], lam[qidx(i, j, 2)], lT[i][j][0],
						lT[i][j][1], lT[i][j][2], lvz);
#pragma unroll
		for (int i = 0; i < 3; i++)
#pragma unroll
			for (int k = 0; k < 3; k++)
				bwd_adj(U[i][0][k], U[i][1][k], vy, lT[i][0][k], lT[i][1][k], lT[i][2][k], lU[i][0][k], lU[i][1][k], lU[i][2][k], lvy);
#pragma unroll
		for (int j = 0; j < 3; j++)
#pragma unroll
			for (int k = 0; k < 3; k++)
				bwd_adj(K[j][k][0], K[j][k][1], vx, lU[0][j][k], lU[1][j][k], lU[2][j][k], lK[j][k][0], lK[j][k][1], lK[j][k][2], lvx);
		// (the first-order inputs hx, hy, hz are constants: post-collision
		// k_100 = -(j_x - m0 u_x) = F_x / 2 identically)
		real lm0 = lK[0][0][0];
		real la = lK[0][0][2], lb = lK[2][0][0], lc = lK[0][2][0];
		real ls = lK[1][0][1], lp = lK[0][1][1], lqq = lK[1][1][0];
		real lks220 = lK[2][0][2], lks211 = lK[1][1][2], lks121 = lK[2][1][1], lks202 = lK[0][2][2], lks112 = lK[1][2][1];
		real lks022 = lK[2][2][0];
		const real lks222 = lK[2][2][2];
		real lrk = 0;
		// ks222 = Y1 rk - Y2 rk^2
		const real lY1 = lks222 * rk, lY2 = -lks222 * rk * rk;
		lrk += lks222 * (Y1 - (real) 2 * Y2 * rk);
		la += lY1 * ks022; lks022 += lY1 * a;
		lb += lY1 * ks202; lks202 += lY1 * b;
		lc += lY1 * ks220; lks220 += lY1 * c;
		lqq += (real) 4 * lY1 * ks211; lks211 += (real) 4 * lY1 * qq;
		lp += (real) 4 * lY1 * ks121;  lks121 += (real) 4 * lY1 * p;
		ls += (real) 4 * lY1 * ks112;  lks112 += (real) 4 * lY1 * s;
		ls += lY2 * ((real) 16 * p * qq + (real) 8 * s * c);
		lp += lY2 * ((real) 16 * s * qq + (real) 8 * p * b);
		lqq += lY2 * ((real) 16 * s * p + (real) 8 * qq * a);
		lb += lY2 * ((real) 4 * p * p + (real) 2 * a * c);
		la += lY2 * ((real) 4 * qq * qq + (real) 2 * b * c);
		lc += lY2 * ((real) 4 * s * s + (real) 2 * a * b);
		// ks_xyz = X_xyz rk
		lrk += lks211 * X211 + lks121 * X121 + lks112 * X112 + lks220 * X220 + lks022 * X022 + lks202 * X202;
		const real l211 = lks211 * rk, l121 = lks121 * rk, l112 = lks112 * rk;
		const real l220 = lks220 * rk, l022 = lks022 * rk, l202 = lks202 * rk;
		la += l211 * qq; lqq += l211 * a; lp += (real) 2 * l211 * s; ls += (real) 2 * l211 * p;
		lb += l121 * p; lp += l121 * b; ls += (real) 2 * l121 * qq; lqq += (real) 2 * l121 * s;
		lc += l112 * s; ls += l112 * c; lqq += (real) 2 * l112 * p; lp += (real) 2 * l112 * qq;
		lb += l220 * a; la += l220 * b; ls += (real) 4 * l220 * s;
		lc += l022 * b; lb += l022 * c; lqq += (real) 4 * l022 * qq;
		la += l202 * c; lc += l202 * a; lp += (real) 4 * l202 * p;
		// a, b, c <- E33, E34, E35 = m0
		const real lE33 = third * (la - (real) 2 * lb + lc);
		const real lE34 = third * (la + lb - (real) 2 * lc);
		lm0 += third * (la + lb + lc);
		const real lk200 = om * (lE33 + lE34), lk020 = -om * lE33, lk002 = -om * lE34;
		const real lk110 = om * ls, lk101 = om * lp, lk011 = om * lqq;
		if (! rho_fixed)
			lm0 -= lrk * rk * rk;
		// central moments <- raw moments, velocity, mass
		real ljx = 0, ljy = 0, ljz = 0;
		lvx += lk200 * ((real) 2 * (m0 * vx - jx)); ljx -= (real) 2 * vx * lk200; lm0 += vx * vx * lk200;
		lvy += lk020 * ((real) 2 * (m0 * vy - jy)); ljy -= (real) 2 * vy * lk020; lm0 += vy * vy * lk020;
		lvz += lk002 * ((real) 2 * (m0 * vz - jz)); ljz -= (real) 2 * vz * lk002; lm0 += vz * vz * lk002;
		lvx += lk110 * (m0 * vy - jy); lvy += lk110 * (m0 * vx - jx); ljy -= vx * lk110; ljx -= vy * lk110; lm0 += vx * vy * lk110;
		lvx += lk101 * (m0 * vz - jz); lvz += lk101 * (m0 * vx - jx); ljz -= vx * lk101; ljx -= vz * lk101; lm0 += vx * vz * lk101;
		lvy += lk011 * (m0 * vz - jz); lvz += lk011 * (m0 * vy - jy); ljz -= vy * lk011; ljy -= vz * lk011; lm0 += vy * vz * lk011;
		// u = (j + F/2) / m0
		ljx += lvx * rv; ljy += lvy * rv; ljz += lvz * rv;
		lm0 -= (lvx * vx + lvy * vy + lvz * vz) * rv;
		// raw moments are linear in f
#pragma unroll
		for (int q = 0; q < 27; q++) {
			const real cx = (real) CX[q], cy = (real) CY[q], cz = (real) CZ[q];
			mu[q] = (Mu) (lm0 + cx * ljx + cy * ljy + cz * ljz + (cx * cx) * lk200 + (cy * cy) * lk020 + (cz * cz) * lk002
						  + (cx * cy) * lk110 + (cx * cz) * lk101 + (cy * cz) * lk011);
		}
	}

	// gradient with respect to u of sum_q lam_q feq_q(1, u), feq = D3Q27_EQ_INV_CUM
	template <int c>
	static LBM_DEV real phi(real v)
	{
		if constexpr (c == 0)
			return (real) 3 * v * v - (real) 2;
		else if constexpr (c > 0)
			return (real) 3 * v * v + (real) 3 * v + (real) 1;
		else
			return (real) 3 * v * v - (real) 3 * v + (real) 1;
	}
	template <int c>
	static LBM_DEV real dphi(real v)
	{
		return (real) 6 * v + (real) (3 * c);
	}
	template <int q, typename L>
	static LBM_DEV void eq_grad_q(const L* lam, real vx, real vy, real vz, real& gx, real& gy, real& gz)
	{
		constexpr int nz = (CX[q] != 0) + (CY[q] != 0) + (CZ[q] != 0);
		constexpr double coef = nz == 0 ? -1.0 / 27.0 : nz == 1 ? 1.0 / 54.0 : nz == 2 ? -1.0 / 108.0 : 1.0 / 216.0;
		const real w = (real) coef * (real) lam[q];
		const real px = phi<CX[q]>(vx), py = phi<CY[q]>(vy), pz = phi<CZ[q]>(vz);
		gx += w * dphi<CX[q]>(vx) * py * pz;
		gy += w * px * dphi<CY[q]>(vy) * pz;
		gz += w * px * py * dphi<CZ[q]>(vz);
	}
	template <int q = 0, typename L>
	static LBM_DEV void eq_grad(const L* lam, real vx, real vy, real vz, real& gx, real& gy, real& gz)
	{
		if constexpr (q < 27) {
			eq_grad_q<q>(lam, vx, vy, vz, gx, gy, gz);
			eq_grad<q + 1>(lam, vx, vy, vz, gx, gy, gz);
		}
	}
};

// adjoint arrays of one step
struct AdjPtrs
{
	const void* lam_next;  // adjoint of the state the forward step wrote (27 XYZ)
	void* lam;             // adjoint of the state the forward step read (27 XYZ, output)
	double* grad;          // d/d(inflow velocity), 3 x (Y Z) indexed Y z + y, accumulated; or null
	const void* tape;      // raw moments recorded by the forward step (tape_index layout), or null
};

// One thread per site.  Pass SPECIAL = false writes, for every site and q, the
// site's contribution to the slot it would pull from under periodic wrapping
// (that map is a bijection, so every output slot is written exactly once:
// plain stores, no clearing); slots that the site really pulls from somewhere
// else (array-edge clamps of non-periodic sites, OUTFLOW_RIGHT's pull from
// x-1, NOTHING sites that pull nothing) get 0 there.  Pass SPECIAL = true then
// adds those contributions to the slots actually pulled, atomically.
// The faces of the box [0,X)x[0,Y)x[0,Z), each site once: x faces whole,
// then the y faces of the interior x, then the z faces of the interior x and
// y.  Only face sites can pull from a source other than the periodic wrap
// (non-periodic neighbour indices clamp at the array edge), and OUTFLOW_RIGHT
// belongs on the x = X-1 face.
inline uint64_t shell_count(int64_t X, int64_t Y, int64_t Z)
{
	const uint64_t fx = X > 1 ? 2 : 1, fy = Y > 1 ? 2 : 1, fz = Z > 1 ? 2 : 1;
	const uint64_t ix = X > 2 ? X - 2 : 0, iy = Y > 2 ? Y - 2 : 0;
	return fx * Y * Z + ix * fy * Z + ix * iy * fz;
}

LBM_DEV bool shell_coords(const StepArgs& A, int64_t& x, int64_t& z, int64_t& y)
{
	uint64_t i = (uint64_t) blockIdx.x * blockDim.x + threadIdx.x;
	const uint64_t X = A.X, Y = A.Y, Z = A.Z;
	const uint64_t fx = X > 1 ? 2 : 1, fy = Y > 1 ? 2 : 1, fz = Z > 1 ? 2 : 1;
	const uint64_t ix = X > 2 ? X - 2 : 0, iy = Y > 2 ? Y - 2 : 0;
	const uint64_t n1 = fx * Y * Z, n2 = ix * fy * Z, n3 = ix * iy * fz;
	if (i < n1) {
		const uint64_t r = i % (Y * Z);
		x = i / (Y * Z) ? (int64_t) X - 1 : 0;
		z = (int64_t) (r / Y);
		y = (int64_t) (r % Y);
		return true;
	}
	i -= n1;
	if (i < n2) {
		const uint64_t r = i % (fy * Z);
		x = 1 + (int64_t) (i / (fy * Z));
		z = (int64_t) (r / fy);
		y = r % fy ? (int64_t) Y - 1 : 0;
		return true;
	}
	i -= n2;
	if (i < n3) {
		const uint64_t r = i % (iy * fz);
		x = 1 + (int64_t) (i / (iy * fz));
		y = 1 + (int64_t) (r / fz);
		z = r % fz ? (int64_t) Z - 1 : 0;
		return true;
	}
	return false;
}

// SPECIAL = false: every site, plain stores to the slots it owns under the
// periodic wrap (0 for slots whose true reader is elsewhere).  SPECIAL = true:
// the face sites only, atomic adds of those contributions to their true
// sources, after the first pass.
// (256, 2): two waves per SIMD; the FP64 form then spills a few registers
// instead of running one wave per SIMD at 264
template <int ST, bool SPECIAL, int TAPE = 0, bool O32 = false>
__global__ void __launch_bounds__(256, 2) k_adjoint_ab(const StepArgs A, const AdjPtrs P)
{
	using S = Storage<ST>;
	using real = typename S::real;
	using elem = typename S::elem;
	using D = AdjCum<real>;
	// lambda in and mu out in the element type where it is narrower than the
	// arithmetic (float fields with a double sweep): fewer registers
	using lreal = std::conditional_t<(sizeof(elem) < sizeof(real)), elem, real>;
	int64_t x, z, y;
	if (SPECIAL ? ! shell_coords(A, x, z, y) : ! block_coords(A, x, z, y))
		return;
	const int64_t s0 = (x * A.Zs + z) * A.Ys + y;
	const int16_t gi = A.map ? A.map[s0] : (int16_t) GEO_PERIODIC;
	const bool per = gi == GEO_PERIODIC;
	const Nb NX = neighbours<false>(x, A.X, 0, false, per);
	const Nb NY = neighbours<false>(y, A.Y, 0, false, per);
	const Nb NZ = neighbours<false>(z, A.Z, 0, false, per);
	Nb PX = NX;
	if (gi == GEO_OUTFLOW_RIGHT)
		PX.p = PX.o = NX.m;
	const Nb WX = neighbours<false>(x, A.X, 0, false, true);
	const Nb WY = neighbours<false>(y, A.Y, 0, false, true);
	const Nb WZ = neighbours<false>(z, A.Z, 0, false, true);
	auto sidx = [&](int64_t xs, int64_t ys, int64_t zs) { return (xs * A.Zs + zs) * A.Ys + ys; };
	uint32_t irregular = 0;
#pragma unroll
	for (int q = 0; q < 27; q++) {
		const bool same = PX[-CX[q]] == WX[-CX[q]] && NY[-CY[q]] == WY[-CY[q]] && NZ[-CZ[q]] == WZ[-CZ[q]];
		if (gi == GEO_NOTHING || ! same)
			irregular |= 1u << q;
	}
	if (SPECIAL && (gi == GEO_NOTHING || irregular == 0))
		return;
	if (SPECIAL && (gi == GEO_WALL || gi == GEO_INFLOW)) {
		// bounce-back: mu_q = lam_next[opp(q)], read only for the irregular directions
		// (the face sites are one per cache line here); inflow: mu = 0
		if (gi == GEO_WALL) {
			const elem* Ln = (const elem*) P.lam_next;
			elem* L = (elem*) P.lam;
#pragma unroll
			for (int q = 0; q < 27; q++)
				if ((irregular >> q) & 1u)
					atomicAdd(L + q * A.XYZ + sidx(PX[-CX[q]], NY[-CY[q]], NZ[-CZ[q]]),
							  (elem) S::ld(Ln + opp(q) * A.XYZ + s0));
		}
		return;
	}
	const bool on_shell = x == 0 || x == A.X - 1 || y == 0 || y == A.Y - 1 || z == 0 || z == A.Z - 1;
	// an irregular pull away from the faces (OUTFLOW_RIGHT inside the box) has no fix-up: poison it
	const bool unsupported = gi != GEO_NOTHING && irregular != 0 && ! on_shell;

	lreal mu[27];
	if (gi == GEO_NOTHING || unsupported) {
		const lreal v = unsupported ? (lreal) __builtin_nan("") : (lreal) 0;
#pragma unroll
		for (int q = 0; q < 27; q++)
			mu[q] = v;
	}
	else {
		const elem* Ln = (const elem*) P.lam_next;
		lreal lam[27];
#pragma unroll
		for (int q = 0; q < 27; q++)
			lam[q] = (lreal) S::ld(slot0<O32, false>(Ln, q, A, s0));
		// raw moments of the pulled populations: from the state, or recorded by the forward step
		const bool needs_m = gi == GEO_FLUID || gi == GEO_PERIODIC || gi == GEO_OUTFLOW_RIGHT || gi == GEO_OUTFLOW_EQ;
		real m[D::NMOM];
#pragma unroll
		for (int k = 0; k < D::NMOM; k++)
			m[k] = 0;
		if (needs_m) {
			if constexpr (TAPE == 0) {
				const elem* F = (const elem*) A.df0;
				real f[27];
#pragma unroll
				for (int q = 0; q < 27; q++)
					f[q] = S::ld(slot<O32, false>(F, q, A, PX[-CX[q]], NY[-CY[q]], NZ[-CZ[q]]));
				D::raw_moments(f, m);
			}
			else {
				using TT = std::conditional_t<TAPE == 2, double, float>;
				const TT* tp = (const TT*) P.tape;
#pragma unroll
				for (int k = 0; k < D::NMOM; k++)
					m[k] = (real) tp[tape_index(k, s0)];
			}
		}
		switch (gi) {
			case GEO_FLUID:
			case GEO_PERIODIC:
			case GEO_OUTFLOW_RIGHT:
				D::vjp_m(m, lam, (real) A.fx, (real) A.fy, (real) A.fz, (real) A.omega1, gi == GEO_OUTFLOW_RIGHT, mu);
				break;
			case GEO_WALL:
#pragma unroll
				for (int q = 0; q < 27; q++)
					mu[q] = lam[opp(q)];
				break;
			case GEO_INFLOW:
				{
#pragma unroll
					for (int q = 0; q < 27; q++)
						mu[q] = 0;
					if (! SPECIAL && P.grad) {
						const int64_t iy = A.Y * z + y;
						const real ux = (real) inflow_component(A, 0, iy);
						const real uy = (real) inflow_component(A, 1, iy);
						const real uz = (real) inflow_component(A, 2, iy);
						real gx = 0, gy = 0, gz = 0;
						D::eq_grad(lam, ux, uy, uz, gx, gy, gz);
						const int64_t n = A.Y * A.Z;
						atomicAdd(P.grad + iy, (double) gx);
						atomicAdd(P.grad + n + iy, (double) gy);
						atomicAdd(P.grad + 2 * n + iy, (double) gz);
					}
					break;
				}
			case GEO_OUTFLOW_EQ:
				{
					// f_out = feq(1, u(f_in)), u = (j + F/2) / m0
					const real m0 = m[0], jx = m[1], jy = m[2], jz = m[3];
					const real rv = (real) 1 / m0;
					const real ux = (jx + (real) 0.5 * (real) A.fx) * rv, uy = (jy + (real) 0.5 * (real) A.fy) * rv,
							   uz = (jz + (real) 0.5 * (real) A.fz) * rv;
					real gx = 0, gy = 0, gz = 0;
					D::eq_grad(lam, ux, uy, uz, gx, gy, gz);
#pragma unroll
					for (int q = 0; q < 27; q++)
						mu[q] = (lreal) ((gx * ((real) CX[q] - ux) + gy * ((real) CY[q] - uy) + gz * ((real) CZ[q] - uz)) * rv);
					break;
				}
			default:
				// no adjoint for this geometry code: make the result unusable, visibly
#pragma unroll
				for (int q = 0; q < 27; q++)
					mu[q] = (lreal) __builtin_nan("");
				break;
		}
	}

	elem* L = (elem*) P.lam;
	if (! SPECIAL) {
#pragma unroll
		for (int q = 0; q < 27; q++) {
			const real v = (irregular >> q) & 1u && ! unsupported ? (real) 0 : mu[q];
			S::st(slot<O32, false>(L, q, A, WX[-CX[q]], WY[-CY[q]], WZ[-CZ[q]]), v);
		}
	}
	else {
#pragma unroll
		for (int q = 0; q < 27; q++)
			if ((irregular >> q) & 1u)
				atomicAdd(L + q * A.XYZ + sidx(PX[-CX[q]], NY[-CY[q]], NZ[-CZ[q]]), (elem) mu[q]);
	}
}

// ===========================================================================
// Adjoint of the D3Q7 part of one coupled A-B step (ade_site), for a given
// velocity field: the passive scalar does not act on the flow, and the D3Q7
// collision is linear in the populations for a given u (SRT, MRT and CLBM
// relax moments whose equilibria are linear in phi), so
//   lambda <- C(u)^T lambda_next scattered back along the pull,
//   dJ/dS  += sum_q w_q lambda_next_q at the sites that collide.
// C(u)^T is applied directly (Ade7::collide_adjoint: each operator is the
// identity or a multiple of it plus a few rank-one terms).
// ===========================================================================
struct AdeAdjArgs
{
	const void* u;          // vx, vy, vz planes (XYZ each) of the step's velocity, macro_real
	const void* lam_next;   // 7 XYZ of the compute type
	void* lam;              // 7 XYZ, output
	double* grad_source;    // XYZ, accumulated, or null
	const int16_t* map;     // D3Q7 codes, or null (all periodic)
	const void* diff;       // per-site diffusivity (macro_real) or null
	double D;
	int32_t collision;
};

template <int ST, bool SPECIAL>
__global__ void __launch_bounds__(256) k_ade_adjoint_ab(const StepArgs A, const AdeAdjArgs P)
{
	using K = Ade7<ST>;
	using real = typename K::real;
	int64_t x, z, y;
	if (SPECIAL ? ! shell_coords(A, x, z, y) : ! block_coords(A, x, z, y))
		return;
	const int64_t s0 = (x * A.Zs + z) * A.Ys + y;
	// the pull uses the NSE site's periodicity, as in the coupled kernel
	const int16_t gn = A.map ? A.map[s0] : (int16_t) GEO_PERIODIC;
	const bool per = gn == GEO_PERIODIC;
	const int16_t ga = P.map ? P.map[s0] : (int16_t) A_PERIODIC;
	const Nb NX = neighbours<false>(x, A.X, 0, false, per);
	const Nb NY = neighbours<false>(y, A.Y, 0, false, per);
	const Nb NZ = neighbours<false>(z, A.Z, 0, false, per);
	Nb PX = NX;
	if (ga == A_OUTFLOW_RIGHT)
		PX.p = PX.o = NX.m;
	const Nb WX = neighbours<false>(x, A.X, 0, false, true);
	const Nb WY = neighbours<false>(y, A.Y, 0, false, true);
	const Nb WZ = neighbours<false>(z, A.Z, 0, false, true);
	auto sidx = [&](int64_t xs, int64_t ys, int64_t zs) { return (xs * A.Zs + zs) * A.Ys + ys; };
	uint32_t irregular = 0;
#pragma unroll
	for (int q = 0; q < 7; q++) {
		const bool same = PX[-CX[q]] == WX[-CX[q]] && NY[-CY[q]] == WY[-CY[q]] && NZ[-CZ[q]] == WZ[-CZ[q]];
		if (ga == A_NOTHING || ! same)
			irregular |= 1u << q;
	}
	if (SPECIAL && (ga == A_NOTHING || irregular == 0))
		return;
	const real* Ln = (const real*) P.lam_next;
	real* L = (real*) P.lam;
	if (SPECIAL && (ga == A_WALL || ga == A_INFLOW)) {
#pragma unroll
		for (int q = 0; q < 7; q++)
			if (ga == A_WALL && ((irregular >> q) & 1u))
				atomicAdd(L + q * A.XYZ + sidx(PX[-CX[q]], NY[-CY[q]], NZ[-CZ[q]]), Ln[opp(q) * A.XYZ + s0]);
		return;
	}
	const bool on_shell = x == 0 || x == A.X - 1 || y == 0 || y == A.Y - 1 || z == 0 || z == A.Z - 1;
	const bool unsupported = ga != A_NOTHING && irregular != 0 && ! on_shell;
	real mu[7];
	if (ga == A_NOTHING || unsupported) {
		const real v = unsupported ? (real) __builtin_nan("") : (real) 0;
#pragma unroll
		for (int q = 0; q < 7; q++)
			mu[q] = v;
	}
	else {
		real lam[7];
#pragma unroll
		for (int q = 0; q < 7; q++)
			lam[q] = Ln[q * A.XYZ + s0];
		switch (ga) {
			case A_FLUID:
			case A_PERIODIC:
			case A_SOLID:
			case A_OUTFLOW_RIGHT:
				{
					auto macro_at = [&](const void* p, int64_t i) -> real {
						return A.macro_f64 ? (real) ((const double*) p)[i] : (real) ((const float*) p)[i];
					};
					const real vx = macro_at(P.u, s0), vy = macro_at(P.u, A.XYZ + s0), vz = macro_at(P.u, 2 * A.XYZ + s0);
					const real D = P.diff ? macro_at(P.diff, s0) : (real) P.D;
					K::collide_adjoint(P.collision, lam, mu, vx, vy, vz, D);
					if (! SPECIAL && P.grad_source) {
						real g = 0;
#pragma unroll
						for (int q = 0; q < 7; q++)
							g += (real) w7(q) * lam[q];
						P.grad_source[s0] += (double) g;
					}
					break;
				}
			case A_WALL:
#pragma unroll
				for (int q = 0; q < 7; q++)
					mu[q] = lam[opp(q)];
				break;
			case A_INFLOW:
#pragma unroll
				for (int q = 0; q < 7; q++)
					mu[q] = 0;
				break;
			default:
#pragma unroll
				for (int q = 0; q < 7; q++)
					mu[q] = (real) __builtin_nan("");
				break;
		}
	}
	if (! SPECIAL) {
#pragma unroll
		for (int q = 0; q < 7; q++) {
			const real v = (irregular >> q) & 1u && ! unsupported ? (real) 0 : mu[q];
			L[q * A.XYZ + sidx(WX[-CX[q]], WY[-CY[q]], WZ[-CZ[q]])] = v;
		}
	}
	else {
#pragma unroll
		for (int q = 0; q < 7; q++)
			if ((irregular >> q) & 1u)
				atomicAdd(L + q * A.XYZ + sidx(PX[-CX[q]], NY[-CY[q]], NZ[-CZ[q]]), mu[q]);
	}
}

// the D3Q7 adjoint step (built in lbm_adjoint.hip)
int launch_ade_adjoint(const tnl_lbm_config* cfg, const StepArgs& A, const AdeAdjArgs& P, unsigned threads,
					   unsigned grid, hipStream_t st);

// both passes of one adjoint step (built in lbm_adjoint.hip); tape_kind 0:
// Jacobian from the state df0, 1: from a float tape, 2: from a double tape
int launch_adjoint(const tnl_lbm_config* cfg, const StepArgs& A, const AdjPtrs& P, int tape_kind, unsigned threads,
				   unsigned grid, hipStream_t st);
// one A-B step of the default cumulant that records the raw moments in A.tape
int launch_record(const tnl_lbm_config* cfg, const StepArgs& A, int tape_kind, unsigned threads, unsigned grid,
				  hipStream_t st);

}  // namespace lbm
