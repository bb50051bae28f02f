// Launches of the adjoint step kernels (absolute storage forms, A-B pattern,
// default cumulant); a translation unit of its own so that it compiles in
// parallel with the forward kernels.
#include "lbm_adjoint.hpp"

namespace lbm {

namespace {

template <int ST, int TAPE>
int adjoint_for(const StepArgs& A0, const AdjPtrs& P, unsigned threads, unsigned grid, hipStream_t st)
{
	StepArgs A = A0;
	linearize(A, threads, grid);
	// FP64: 32-bit slot offsets where every population array (state, lambda
	// in and out) holds at most 4 GiB: 2 spilled registers instead of 4 and
	// -68 VALU, ~2% faster on the C4 duct; FP32 measured ~2% slower with them
	constexpr bool F64MATH = std::is_same<typename Storage<ST>::real, double>::value;
	bool o32 = false;
	if constexpr (F64MATH)
		o32 = narrow_slots(A.XYZ, sizeof(typename Storage<ST>::elem));
	if (o32) {
		if constexpr (F64MATH)
			hipLaunchKernelGGL((k_adjoint_ab<ST, false, TAPE, true>), dim3(grid), dim3(threads), 0, st, A, P);
	}
	else
		hipLaunchKernelGGL((k_adjoint_ab<ST, false, TAPE>), dim3(grid), dim3(threads), 0, st, A, P);
	if (int rc = hip_check(hipGetLastError(), "adjoint launch"))
		return rc;
	const uint64_t shell = shell_count(A.X, A.Y, A.Z);
	const unsigned sgrid = (unsigned) ((shell + 255) / 256);
	hipLaunchKernelGGL((k_adjoint_ab<ST, true, TAPE>), dim3(sgrid), dim3(256), 0, st, A, P);
	return hip_check(hipGetLastError(), "adjoint fix-up launch");
}

template <int ST, int TAPE>
int record_for(const StepArgs& A0, unsigned threads, unsigned grid, hipStream_t st)
{
	// (no row lists: the F64 and F32 recording steps on the C4 duct measured
	// 1-4% slower with them, the extra launch of its inflow/outflow rows)
	StepArgs A = A0;
	linearize(A, threads, grid);
	// 32-bit slot offsets where each population array holds at most 4 GiB:
	// -124 (FP64) / -151 (FP32) VALU, ~4% on the C4 duct (kernel trace)
	if (narrow_slots(A.XYZ, sizeof(typename Storage<ST>::elem)))
		hipLaunchKernelGGL((k_general<ST, TNL_LBM_COLL_CUM, TNL_LBM_PATTERN_AB, TAPE, true>), dim3(grid), dim3(threads), 0, st, A);
	else
		hipLaunchKernelGGL((k_general<ST, TNL_LBM_COLL_CUM, TNL_LBM_PATTERN_AB, TAPE>), dim3(grid), dim3(threads), 0, st, A);
	return hip_check(hipGetLastError(), "recording step launch");
}

template <int ST>
int ade_adjoint_for(const StepArgs& A0, const AdeAdjArgs& P, unsigned threads, unsigned grid, hipStream_t st)
{
	StepArgs A = A0;
	linearize(A, threads, grid);
	hipLaunchKernelGGL((k_ade_adjoint_ab<ST, false>), dim3(grid), dim3(threads), 0, st, A, P);
	if (int rc = hip_check(hipGetLastError(), "D3Q7 adjoint launch"))
		return rc;
	const uint64_t shell = shell_count(A.X, A.Y, A.Z);
	hipLaunchKernelGGL((k_ade_adjoint_ab<ST, true>), dim3((unsigned) ((shell + 255) / 256)), dim3(256), 0, st, A, P);
	return hip_check(hipGetLastError(), "D3Q7 adjoint fix-up launch");
}

}  // namespace

int launch_ade_adjoint(const tnl_lbm_config* cfg, const StepArgs& A, const AdeAdjArgs& P, unsigned threads,
					   unsigned grid, hipStream_t st)
{
	// the compute type follows the storage: double for F64, float for the others
	if (cfg->storage == TNL_LBM_STORE_F64)
		return ade_adjoint_for<TNL_LBM_STORE_F64>(A, P, threads, grid, st);
	return ade_adjoint_for<TNL_LBM_STORE_F32>(A, P, threads, grid, st);
}

int launch_adjoint(const tnl_lbm_config* cfg, const StepArgs& A, const AdjPtrs& P, int tape_kind, unsigned threads,
				   unsigned grid, hipStream_t st)
{
	if (cfg->storage == TNL_LBM_STORE_F64) {
		if (tape_kind == 1)
			return adjoint_for<TNL_LBM_STORE_F64, 1>(A, P, threads, grid, st);
		if (tape_kind == 2)
			return adjoint_for<TNL_LBM_STORE_F64, 2>(A, P, threads, grid, st);
		return adjoint_for<TNL_LBM_STORE_F64, 0>(A, P, threads, grid, st);
	}
	// F32 storage: float fields and recordings, double arithmetic.  In float
	// the sweep's rounding (a few ulp per step) accumulates to ~1e-6 of the
	// gradient over 2000 steps; in double it stays below 1e-7, for ~10% more
	// adjoint-step time.
	if (tape_kind == 1)
		return adjoint_for<STORE_F32_F64MATH, 1>(A, P, threads, grid, st);
	return adjoint_for<STORE_F32_F64MATH, 0>(A, P, threads, grid, st);
}

int launch_record(const tnl_lbm_config* cfg, const StepArgs& A, int tape_kind, unsigned threads, unsigned grid,
				  hipStream_t st)
{
	if (cfg->storage == TNL_LBM_STORE_F64)
		return tape_kind == 2 ? record_for<TNL_LBM_STORE_F64, 2>(A, threads, grid, st)
							  : record_for<TNL_LBM_STORE_F64, 1>(A, threads, grid, st);
	return record_for<TNL_LBM_STORE_F32, 1>(A, threads, grid, st);
}

}  // namespace lbm
