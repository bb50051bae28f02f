// Launches of the adjoint step kernels (absolute storage forms, A-B pattern,
// default cumulant); a translation unit of its own so that it compiles in
// parallel with the forward kernels.
#include "lbm_adjoint.hpp"

namespace lbm {

namespace {

template <int ST>
int adjoint_for(const StepArgs& A0, const AdjPtrs& P, unsigned threads, unsigned grid, hipStream_t st)
{
	StepArgs A = A0;
	linearize(A, threads, grid);
	hipLaunchKernelGGL((k_adjoint_ab<ST, false>), dim3(grid), dim3(threads), 0, st, A, P);
	if (int rc = hip_check(hipGetLastError(), "adjoint launch"))
		return rc;
	const uint64_t shell = shell_count(A.X, A.Y, A.Z);
	const unsigned sgrid = (unsigned) ((shell + 255) / 256);
	hipLaunchKernelGGL((k_adjoint_ab<ST, true>), dim3(sgrid), dim3(256), 0, st, A, P);
	return hip_check(hipGetLastError(), "adjoint fix-up launch");
}

}  // namespace

int launch_adjoint(const tnl_lbm_config* cfg, const StepArgs& A, const AdjPtrs& P, unsigned threads, unsigned grid,
				   hipStream_t st)
{
	if (cfg->storage == TNL_LBM_STORE_F64)
		return adjoint_for<TNL_LBM_STORE_F64>(A, P, threads, grid, st);
	return adjoint_for<TNL_LBM_STORE_F32>(A, P, threads, grid, st);
}

}  // namespace lbm
