// Step launches of the alternative D3Q27 operators: MRT_LES (col_mrt.h) and
// the entropic KBC family N1-N4 / C1-C4 (col_kbc_n.h, col_kbc_c.h), for the
// absolute storage forms (F64, F32) as in the reference.  A translation unit
// of its own so that the library's kernels compile in parallel.
#include "lbm_step.hpp"

namespace lbm {

namespace {

template <int ST>
int alt_for(const tnl_lbm_config* cfg, const StepArgs& A, unsigned threads, unsigned grid, hipStream_t st)
{
	switch (cfg->collision) {
		case TNL_LBM_COLL_MRT_LES: return launch_step<ST, TNL_LBM_COLL_MRT_LES>(cfg, A, threads, grid, st);
		case TNL_LBM_COLL_KBC_N1: return launch_step<ST, TNL_LBM_COLL_KBC_N1>(cfg, A, threads, grid, st);
		case TNL_LBM_COLL_KBC_N2: return launch_step<ST, TNL_LBM_COLL_KBC_N2>(cfg, A, threads, grid, st);
		case TNL_LBM_COLL_KBC_N3: return launch_step<ST, TNL_LBM_COLL_KBC_N3>(cfg, A, threads, grid, st);
		case TNL_LBM_COLL_KBC_N4: return launch_step<ST, TNL_LBM_COLL_KBC_N4>(cfg, A, threads, grid, st);
		case TNL_LBM_COLL_KBC_C1: return launch_step<ST, TNL_LBM_COLL_KBC_C1>(cfg, A, threads, grid, st);
		case TNL_LBM_COLL_KBC_C2: return launch_step<ST, TNL_LBM_COLL_KBC_C2>(cfg, A, threads, grid, st);
		case TNL_LBM_COLL_KBC_C3: return launch_step<ST, TNL_LBM_COLL_KBC_C3>(cfg, A, threads, grid, st);
		default: return launch_step<ST, TNL_LBM_COLL_KBC_C4>(cfg, A, threads, grid, st);
	}
}

}  // namespace

int launch_alt(const tnl_lbm_config* cfg, const StepArgs& A, unsigned threads, unsigned grid, hipStream_t st)
{
	// the callers have validated the combination: absolute storage, collision 3..11
	if (cfg->storage == TNL_LBM_STORE_F64)
		return alt_for<TNL_LBM_STORE_F64>(cfg, A, threads, grid, st);
	return alt_for<TNL_LBM_STORE_F32>(cfg, A, threads, grid, st);
}

}  // namespace lbm
