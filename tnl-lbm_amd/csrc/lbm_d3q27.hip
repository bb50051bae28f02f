// gfx950 kernels of the D3Q27 lattice update and the C ABI that drives them.
//
// Replaces cudaLBMKernel<NSE> (TNL-LBM include/lbm3d/kernels.h:59-112) and
// the device loops of LBM_BLOCK (lbm_block.hpp:248-305).  Design:
//  * one lattice site per lane, 64-lane waves laid along y (the contiguous
//    axis), so each of the 27 loads and 27 stores of a wave is one
//    contiguous 64-element segment of one population plane;
//  * a block spans one (x, z) row segment, so the row offsets of the site
//    and of its x/z neighbours are wave-uniform and live in scalar
//    registers; only the y index (and its +-1 neighbours) is per lane;
//  * the 1-D grid is remapped so that consecutive row segments land on the
//    same XCD (workgroups are dispatched round-robin over the 8 XCDs), which
//    keeps the cache lines shared by y-neighbouring waves in one L2;
//  * the pure-periodic A-A path reads no geometry map at all; the general
//    path reads the int16 map and applies D3Q27_BC_All (d3q27/bc.h:61-422).
#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <type_traits>

#include "lbm_adjoint.hpp"
#include "lbm_step.hpp"

using namespace lbm;

namespace lbm {

thread_local std::string g_err;
int fail(int code, const std::string& msg)
{
	g_err = msg;
	return code;
}

int hip_check(hipError_t e, const char* what)
{
	if (e != hipSuccess)
		return fail(TNL_LBM_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
	return TNL_LBM_OK;
}

static std::atomic<int> g_wide_addressing{0};
bool narrow_slots_allowed()
{
	return g_wide_addressing.load(std::memory_order_relaxed) == 0;
}

}  // namespace lbm

namespace {

int validate(const tnl_lbm_config* cfg, const tnl_lbm_d3q27_data* SD)
{
	if (! cfg || ! SD)
		return fail(TNL_LBM_ERR_NULL, "null config or data pointer");
	if (cfg->struct_size != sizeof(tnl_lbm_config) || cfg->version != TNL_LBM_ABI_VERSION)
		return fail(TNL_LBM_ERR_VERSION, "tnl_lbm_config struct_size/version mismatch");
	if (SD->struct_size != sizeof(tnl_lbm_d3q27_data) || SD->version != TNL_LBM_ABI_VERSION)
		return fail(TNL_LBM_ERR_VERSION, "tnl_lbm_d3q27_data struct_size/version mismatch");
	if (SD->X <= 0 || SD->Y <= 0 || SD->Z <= 0 || SD->ox < 0 || SD->oy < 0 || SD->oz < 0)
		return fail(TNL_LBM_ERR_DIMS, "lattice sizes must be positive and overlaps non-negative");
	const int64_t stored = (SD->X + 2 * SD->ox) * (SD->Y + 2 * SD->oy) * (SD->Z + 2 * SD->oz);
	if (SD->XYZ < stored)
		return fail(TNL_LBM_ERR_DIMS, "XYZ is smaller than the stored lattice");
	if (SD->Y + 2 * SD->oy > (int64_t) 1 << 30)
		return fail(TNL_LBM_ERR_DIMS, "Y extent too large");
	if (cfg->storage < TNL_LBM_STORE_F64 || cfg->storage > TNL_LBM_STORE_I16_DEV)
		return fail(TNL_LBM_ERR_ARG, "unknown storage form");
	if (cfg->storage == TNL_LBM_STORE_I16_DEV)
		for (int n = 0; n < 4; n++) {
			if (! (SD->store_range_class[n] >= 0))
				return fail(TNL_LBM_ERR_ARG, "store_range_class must not be negative");
			if (! (SD->store_range_class[n] > 0) && ! (SD->store_range > 0))
				return fail(TNL_LBM_ERR_ARG, "store_range (or every store_range_class) must be positive for the fixed-point form");
		}
	if (cfg->pattern != TNL_LBM_PATTERN_AA && cfg->pattern != TNL_LBM_PATTERN_AB)
		return fail(TNL_LBM_ERR_ARG, "unknown streaming pattern");
	if (cfg->macro_real != TNL_LBM_REAL_F64 && cfg->macro_real != TNL_LBM_REAL_F32)
		return fail(TNL_LBM_ERR_ARG, "unknown macro element type");
	if (cfg->collision < TNL_LBM_COLL_CUM || cfg->collision > TNL_LBM_COLL_KBC_C4)
		return fail(TNL_LBM_ERR_UNSUPPORTED, "unknown collision operator");
	if (cfg->storage >= TNL_LBM_STORE_F32_DEV && cfg->collision != TNL_LBM_COLL_CUM)
		return fail(TNL_LBM_ERR_UNSUPPORTED, "deviation storage is built for the default cumulant only");
	if (! SD->df[0] || (cfg->pattern == TNL_LBM_PATTERN_AB && ! SD->df[1]))
		return fail(TNL_LBM_ERR_NULL, "population array missing");
	return TNL_LBM_OK;
}


// omega1 = 1/(3 nu + 1/2) in the arithmetic of the storage form (col_cum.h:195)
double omega1_of(int storage, double nu)
{
	if (storage == TNL_LBM_STORE_F64)
		return 1.0 / (3.0 * nu + 0.5);  // FP64 arithmetic
	volatile float n = (float) nu;
	volatile float t = 3.0f * n;
	volatile float d = t + 0.5f;
	return (double) (1.0f / d);
}


template <int ST>
int dispatch_coll(const tnl_lbm_config* cfg, const StepArgs& A, unsigned threads, unsigned grid, hipStream_t st)
{
	if constexpr (Storage<ST>::dev)
		return launch_step<ST, TNL_LBM_COLL_CUM>(cfg, A, threads, grid, st);
	else {
		switch (cfg->collision) {
			case TNL_LBM_COLL_CUM: return launch_step<ST, TNL_LBM_COLL_CUM>(cfg, A, threads, grid, st);
			case TNL_LBM_COLL_CUM_2017: return launch_step<ST, TNL_LBM_COLL_CUM_2017>(cfg, A, threads, grid, st);
			case TNL_LBM_COLL_CUM_2017_ANTIALIAS: return launch_step<ST, TNL_LBM_COLL_CUM_2017_ANTIALIAS>(cfg, A, threads, grid, st);
			default: return launch_alt(cfg, A, threads, grid, st);
		}
	}
}

#define STORAGE_SWITCH(storage, EXPR)                                                        \
	switch (storage) {                                                                       \
		case TNL_LBM_STORE_F64: { constexpr int ST = TNL_LBM_STORE_F64; return EXPR; }           \
		case TNL_LBM_STORE_F32: { constexpr int ST = TNL_LBM_STORE_F32; return EXPR; }           \
		case TNL_LBM_STORE_F32_DEV: { constexpr int ST = TNL_LBM_STORE_F32_DEV; return EXPR; }   \
		case TNL_LBM_STORE_F16_DEV: { constexpr int ST = TNL_LBM_STORE_F16_DEV; return EXPR; }   \
		case TNL_LBM_STORE_BF16_DEV: { constexpr int ST = TNL_LBM_STORE_BF16_DEV; return EXPR; } \
		default: { constexpr int ST = TNL_LBM_STORE_I16_DEV; return EXPR; }                     \
	}

// fixed-point scale factors (only the I16 form uses them)
// fixed-point bounds per direction class (store_range_class, else store_range)
FixedRange range_of(const tnl_lbm_d3q27_data* SD)
{
	FixedRange r;
	for (int n = 0; n < 4; n++) {
		const double R = SD->store_range_class[n] > 0 ? SD->store_range_class[n] : SD->store_range;
		r.rs[n] = (float) (R / 32767.0);
		r.invr[n] = R > 0 ? (float) (1.0 / R) : 0.0f;
	}
	return r;
}
void set_range(const tnl_lbm_d3q27_data* SD, float* rs, float* invr)
{
	const FixedRange r = range_of(SD);
	for (int n = 0; n < 4; n++) {
		rs[n] = r.rs[n];
		invr[n] = r.invr[n];
	}
}
const int ELEM_BYTES[6] = {8, 4, 4, 2, 2, 2};

int64_t stored_sites(const tnl_lbm_d3q27_data* SD)
{
	return (SD->X + 2 * SD->ox) * (SD->Y + 2 * SD->oy) * (SD->Z + 2 * SD->oz);
}

MacroArgs macro_args(const tnl_lbm_config* cfg, const tnl_lbm_d3q27_data* SD, const void* df)
{
	MacroArgs M;
	M.df = df;
	M.macro = SD->macro;
	M.macro_f64 = cfg->macro_real == TNL_LBM_REAL_F64;
	M.XYZ = SD->XYZ;
	M.X = SD->X; M.Y = SD->Y; M.Z = SD->Z;
	M.ox = SD->ox; M.oy = SD->oy; M.oz = SD->oz;
	M.Ys = SD->Y + 2 * SD->oy;
	M.Zs = SD->Z + 2 * SD->oz;
	set_range(SD, M.rs, M.invr);
	return M;
}

}  // namespace

// Scratch of the allocation-free reduction: the partial sums of its
// RED_BLOCKS workgroups (48 KiB), one copy per GPU, held until the process
// ends.  Host code never passes `scratch` itself (that name is a host-side
// stand-in); it passes the device address hipGetSymbolAddress returns, kept
// per device in g_scratch_addr.
constexpr int SCRATCH_ROWS = RED_BLOCKS;
static __device__ __align__(256) double scratch[SCRATCH_ROWS * RED_N];
constexpr int SCRATCH_MAX_DEVICES = 64;
static std::mutex g_scratch_mu;
static double* g_scratch_addr[SCRATCH_MAX_DEVICES] = {};

// this device's address of `scratch`, looked up on the first call per device
static int scratch_address(double** out)
{
	int dev = 0;
	if (int rc = hip_check(hipGetDevice(&dev), "hipGetDevice"))
		return rc;
	if (dev < 0 || dev >= SCRATCH_MAX_DEVICES)
		return fail(TNL_LBM_ERR_ARG, "device ordinal beyond the scratch table");
	std::lock_guard<std::mutex> lock(g_scratch_mu);
	if (! g_scratch_addr[dev]) {
		void* p = nullptr;
		if (int rc = hip_check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)"))
			return rc;
		if (! p)
			return fail(TNL_LBM_ERR_HIP, "hipGetSymbolAddress(scratch) returned a null address");
		g_scratch_addr[dev] = static_cast<double*>(p);
	}
	*out = g_scratch_addr[dev];
	return TNL_LBM_OK;
}

// ===========================================================================
extern "C" {

uint32_t tnl_lbm_abi_version(void)
{
	return TNL_LBM_ABI_VERSION;
}

const char* tnl_lbm_last_error(void)
{
	return g_err.c_str();
}

int tnl_lbm_set_wide_addressing(int on)
{
	return g_wide_addressing.exchange(on ? 1 : 0);
}

int tnl_lbm_classify_rows(const tnl_lbm_d3q27_data* nse, const tnl_lbm_d3q7_data* ade, uint32_t* rows,
						  int64_t row_count[3], void* stream)
{
	if (! nse || ! rows || ! row_count)
		return fail(TNL_LBM_ERR_NULL, "null data, rows or row_count");
	if (nse->struct_size != sizeof(tnl_lbm_d3q27_data) || nse->version != TNL_LBM_ABI_VERSION)
		return fail(TNL_LBM_ERR_VERSION, "tnl_lbm_d3q27_data struct_size/version mismatch");
	if (ade) {
		if (ade->struct_size != sizeof(tnl_lbm_d3q7_data) || ade->version != TNL_LBM_ABI_VERSION)
			return fail(TNL_LBM_ERR_VERSION, "tnl_lbm_d3q7_data struct_size/version mismatch");
		if (ade->X != nse->X || ade->Y != nse->Y || ade->Z != nse->Z || ade->ox != nse->ox || ade->oy != nse->oy
			|| ade->oz != nse->oz || ade->XYZ != nse->XYZ)
			return fail(TNL_LBM_ERR_DIMS, "NSE and ADE lattices must have the same geometry");
	}
	if (nse->X < 0 || nse->Y < 0 || nse->Z < 0 || nse->X >= 65536 || nse->Z >= 65536)
		return fail(TNL_LBM_ERR_DIMS, "row lists need 0 <= X, Z < 65536");
	for (int c = 0; c < 3; c++)
		row_count[c] = 0;
	const int64_t n = nse->X * nse->Z;
	if (n == 0)
		return TNL_LBM_OK;
	hipStream_t st = (hipStream_t) stream;
	void* buf = nullptr;
	const size_t cls_bytes = ((size_t) n + 255) / 256 * 256;
	if (int rc = hip_check(hipMalloc(&buf, cls_bytes + 3 * sizeof(uint32_t)), "classify scratch"))
		return rc;
	uint8_t* cls = (uint8_t*) buf;
	uint32_t* counts = (uint32_t*) ((char*) buf + cls_bytes);
	const int64_t Ys = nse->Y + 2 * nse->oy, Zs = nse->Z + 2 * nse->oz;
	hipLaunchKernelGGL(k_classify_rows, dim3((unsigned) ((n * 64 + 255) / 256)), dim3(256), 0, st, nse->map,
					   ade ? ade->map : nullptr, nse->X, nse->Y, nse->Z, nse->ox, nse->oy, nse->oz, Ys, Zs, cls);
	hipLaunchKernelGGL(k_row_lists, dim3(1), dim3(1024), 0, st, cls, (uint32_t) n, (uint32_t) nse->Z, rows, counts);
	uint32_t host[3] = {0, 0, 0};
	int rc = hip_check(hipGetLastError(), "classify launch");
	if (! rc)
		rc = hip_check(hipMemcpyAsync(host, counts, sizeof host, hipMemcpyDeviceToHost, st), "classify counts");
	if (! rc)
		rc = hip_check(hipStreamSynchronize(st), "classify");
	const int frc = hip_check(hipFree(buf), "classify scratch free");
	if (rc)
		return rc;
	for (int c = 0; c < 3; c++)
		row_count[c] = host[c];
	return frc;
}

static int make_step_args(const tnl_lbm_config* cfg, const tnl_lbm_d3q27_data* SD, const int64_t offset[3],
						  const int64_t end[3], uint32_t flags, StepArgs& A, unsigned& threads, unsigned& grid);

int tnl_lbm_d3q27_step(const tnl_lbm_config* cfg, const tnl_lbm_d3q27_data* SD, const int64_t offset[3],
					   const int64_t end[3], uint32_t flags, void* stream)
{
	StepArgs A;
	unsigned threads = 0, grid = 0;
	if (int rc = make_step_args(cfg, SD, offset, end, flags, A, threads, grid))
		return rc;
	if (grid == 0)
		return TNL_LBM_OK;
	hipStream_t st = (hipStream_t) stream;
	STORAGE_SWITCH(cfg->storage, (dispatch_coll<ST>(cfg, A, threads, grid, st)));
}

int tnl_lbm_d3q27_stream_only(const tnl_lbm_config* cfg, const tnl_lbm_d3q27_data* SD, void* stream)
{
	if (! cfg || ! SD)
		return fail(TNL_LBM_ERR_NULL, "null config or data pointer");
	if (cfg->pattern != TNL_LBM_PATTERN_AA || SD->map || SD->force_field)
		return fail(TNL_LBM_ERR_UNSUPPORTED, "streaming-only runs the map-free periodic A-A kernels only");
	const int64_t offset[3] = {0, 0, 0};
	const int64_t end[3] = {SD->X, SD->Y, SD->Z};
	StepArgs A;
	unsigned threads = 0, grid = 0;
	if (int rc = make_step_args(cfg, SD, offset, end, 0, A, threads, grid))
		return rc;
	if (grid == 0)
		return TNL_LBM_OK;
	// the fixed-point encode without the fast collision's 2^-n output factor
	for (int n = 0; n < 4; n++)
		A.cls_inv[n] = A.cls_sc[n] > 0 ? (float) (1.0 / ((double) A.cls_sc[n] * 32767.0)) : 0.0f;
	hipStream_t st = (hipStream_t) stream;
	STORAGE_SWITCH(cfg->storage, (launch_periodic<ST, COLL_STREAM_ONLY>(A, threads, grid, st)));
}

// The three row lists of a full-box launch, checked against the box
static int set_lists(StepArgs& A, int64_t X, int64_t Z, const uint32_t* rows, const int64_t count[3])
{
	int64_t sum = 0;
	for (int c = 0; c < 3; c++) {
		if (count[c] < 0)
			return fail(TNL_LBM_ERR_ARG, "negative row count");
		sum += count[c];
	}
	if (sum > X * Z || X >= 65536 || Z >= 65536 || (uint64_t) sum * (uint64_t) A.ny >= (1ull << 31))
		return fail(TNL_LBM_ERR_DIMS, "row lists do not fit the box");
	A.lists = rows;
	for (int c = 0; c < 3; c++)
		A.list_count[c] = count[c];
	return TNL_LBM_OK;
}

static int make_step_args(const tnl_lbm_config* cfg, const tnl_lbm_d3q27_data* SD, const int64_t offset[3],
						  const int64_t end[3], uint32_t flags, StepArgs& A, unsigned& threads, unsigned& grid)
{
	grid = 0;
	if (int rc = validate(cfg, SD))
		return rc;
	if (! offset || ! end)
		return fail(TNL_LBM_ERR_NULL, "offset/end missing");
	const int64_t lo[3] = {SD->X, SD->Y, SD->Z};
	for (int d = 0; d < 3; d++)
		if (offset[d] < 0 || end[d] > lo[d] || end[d] < offset[d])
			return fail(TNL_LBM_ERR_DIMS, "offset/end outside the local lattice");
	if ((flags & TNL_LBM_FLAG_OUTPUT_MACRO) && ! SD->macro)
		return fail(TNL_LBM_ERR_NULL, "output_macro requested without a macro array");
	std::memset(&A, 0, sizeof(A));
	A.df0 = SD->df[0];
	A.df1 = SD->df[1];
	A.macro = SD->macro;
	A.map = SD->map;
	A.XYZ = SD->XYZ;
	A.X = SD->X; A.Y = SD->Y; A.Z = SD->Z;
	A.ox = SD->ox; A.oy = SD->oy; A.oz = SD->oz;
	A.Ys = SD->Y + 2 * SD->oy;
	A.Zs = SD->Z + 2 * SD->oz;
	A.plane_bytes = (uint64_t) SD->XYZ * (uint64_t) ELEM_BYTES[cfg->storage];
	set_range(SD, A.rs, A.invr);
	{
		const bool fast = cfg->storage != TNL_LBM_STORE_F64 && cfg->collision == TNL_LBM_COLL_CUM;
		for (int n = 0; n < 4; n++) {
			double w = 1.0;  // same product order as w27(): one value per class
			for (int d = 0; d < 3; d++)
				w *= d < n ? W1[0] : W1[1];
			A.cls_sc[n] = (float) w * A.rs[n];
			A.cls_inv[n] = (float) (1.0 / w) * A.invr[n] * (fast ? 1.0f / (float) (1 << n) : 1.0f);
		}
	}
	A.offx = (int32_t) offset[0]; A.offy = (int32_t) offset[1]; A.offz = (int32_t) offset[2];
	A.nx = (int32_t) (end[0] - offset[0]); A.ny = (int32_t) (end[1] - offset[1]); A.nz = (int32_t) (end[2] - offset[2]);
	if (A.nx == 0 || A.ny == 0 || A.nz == 0)
		return TNL_LBM_OK;
	threads = A.ny > 128 ? 256 : (A.ny > 64 ? 128 : 64);
	A.nyb = (uint32_t) ((A.ny + threads - 1) / threads);
	const uint64_t nb = (uint64_t) A.nyb * A.nz * A.nx;
	if (nb >= (1ull << 31))
		return fail(TNL_LBM_ERR_DIMS, "lattice box too large for one launch");
	A.nblocks = (uint32_t) nb;
	const uint32_t padded = (A.nblocks + NXCD - 1) / NXCD * NXCD;
	A.per_xcd = padded / NXCD;
	A.even = SD->even_iter ? 1 : 0;
	A.out_macro = (flags & TNL_LBM_FLAG_OUTPUT_MACRO) ? 1 : 0;
	A.macro_f64 = cfg->macro_real == TNL_LBM_REAL_F64;
	A.dist_x = (flags & TNL_LBM_FLAG_DISTRIBUTED_X) != 0;
	A.dist_y = (flags & TNL_LBM_FLAG_DISTRIBUTED_Y) != 0;
	A.dist_z = (flags & TNL_LBM_FLAG_DISTRIBUTED_Z) != 0;
	// a distributed axis reads its neighbours from the overlap layer, which must exist
	if ((A.dist_x && SD->ox < 1) || (A.dist_y && SD->oy < 1) || (A.dist_z && SD->oz < 1))
		return fail(TNL_LBM_ERR_DIMS, "a distributed axis needs an overlap of at least one site");
	if (SD->lbm_viscosity == 0)
		return fail(TNL_LBM_ERR_ARG, "LBM viscosity is 0");  // state.hpp:1240-1245
	A.omega1 = omega1_of(cfg->storage, SD->lbm_viscosity);
	A.nu = SD->lbm_viscosity;
	A.fx = SD->fx; A.fy = SD->fy; A.fz = SD->fz;
	A.inflow[0] = SD->inflow_vx; A.inflow[1] = SD->inflow_vy; A.inflow[2] = SD->inflow_vz;
	if (SD->inflow_profile_real != TNL_LBM_REAL_F64 && SD->inflow_profile_real != TNL_LBM_REAL_F32)
		return fail(TNL_LBM_ERR_ARG, "unknown inflow profile element type");
	for (int k = 0; k < 3; k++)
		A.prof[k] = SD->inflow_profile[k];
	A.prof_f32 = SD->inflow_profile_real == TNL_LBM_REAL_F32;
	A.force = SD->force_field;
	// the row lists apply to a launch over the whole box
	const bool full = offset[0] == 0 && offset[1] == 0 && offset[2] == 0 && end[0] == SD->X && end[1] == SD->Y
					  && end[2] == SD->Z;
	if (full && SD->rows)
		if (int rc = set_lists(A, SD->X, SD->Z, SD->rows, SD->row_count))
			return rc;
	grid = padded;
	return TNL_LBM_OK;
}

static int validate_ade(const tnl_lbm_config* cfg, const tnl_lbm_d3q7_data* ADE)
{
	if (! cfg || ! ADE)
		return fail(TNL_LBM_ERR_NULL, "null config or ADE data");
	if (ADE->struct_size != sizeof(tnl_lbm_d3q7_data) || ADE->version != TNL_LBM_ABI_VERSION)
		return fail(TNL_LBM_ERR_VERSION, "tnl_lbm_d3q7_data struct_size/version mismatch");
	if (! ADE->df[0] || (cfg->pattern == TNL_LBM_PATTERN_AB && ! ADE->df[1]))
		return fail(TNL_LBM_ERR_NULL, "ADE population array missing");
	if (ADE->collision < TNL_LBM_ADE_COLL_SRT || ADE->collision > TNL_LBM_ADE_COLL_CLBM)
		return fail(TNL_LBM_ERR_UNSUPPORTED, "unknown ADE collision");
	if (cfg->storage >= TNL_LBM_STORE_F32_DEV && ADE->store_scale == 0)
		return fail(TNL_LBM_ERR_ARG, "ADE store_scale must be non-zero");
	return TNL_LBM_OK;
}

int tnl_lbm_nse_ade_step(const tnl_lbm_config* cfg, const tnl_lbm_d3q27_data* nse, const tnl_lbm_d3q7_data* ade,
						 const int64_t offset[3], const int64_t end[3], uint32_t flags, void* stream)
{
	StepArgs A;
	unsigned threads = 0, grid = 0;
	if (int rc = make_step_args(cfg, nse, offset, end, flags, A, threads, grid))
		return rc;
	if (int rc = validate_ade(cfg, ade))
		return rc;
	if (ade->X != nse->X || ade->Y != nse->Y || ade->Z != nse->Z || ade->ox != nse->ox || ade->oy != nse->oy
		|| ade->oz != nse->oz || ade->XYZ != nse->XYZ)
		return fail(TNL_LBM_ERR_DIMS, "NSE and ADE lattices must have the same geometry");
	if (cfg->collision != TNL_LBM_COLL_CUM)
		return fail(TNL_LBM_ERR_UNSUPPORTED, "the coupled kernel is built for the default cumulant");
	if ((flags & TNL_LBM_FLAG_OUTPUT_MACRO) && ! ade->macro)
		return fail(TNL_LBM_ERR_NULL, "output_macro requested without an ADE macro array");
	if (grid == 0)
		return TNL_LBM_OK;
	AdeArgs B;
	std::memset(&B, 0, sizeof(B));
	B.df0 = ade->df[0];
	B.df1 = ade->df[1];
	B.macro = ade->macro;
	B.map = ade->map;
	B.diff = ade->diffusion_coefficient;
	B.source = ade->source;
	B.D = ade->lbm_diffusivity;
	B.inflow_phi = ade->inflow_phi;
	B.phi_ref = ade->phi_ref;
	B.scale = ade->store_scale;
	B.collision = ade->collision;
	B.even = ade->even_iter ? 1 : 0;
	hipStream_t st = (hipStream_t) stream;
	// the coupled step takes the lists classified from both maps (the ADE
	// data's), never the NSE lattice's own
	A.lists = nullptr;
	const bool full = offset[0] == 0 && offset[1] == 0 && offset[2] == 0 && end[0] == nse->X && end[1] == nse->Y
					  && end[2] == nse->Z;
	if (full && ade->rows)
		if (int rc = set_lists(A, nse->X, nse->Z, ade->rows, ade->row_count))
			return rc;
	auto go = [&](auto stc) {
		constexpr int ST = decltype(stc)::value;
		// 2-byte forms: 32-bit slot offsets where each population array (D3Q27
		// and D3Q7 share XYZ) holds at most 4 GiB, as in the map kernel
		constexpr bool two = sizeof(typename Storage<ST>::elem) == 2;
		const bool o32 = two && narrow_slots(A.XYZ, 2);
		auto one = [&](auto lc, const StepArgs& C, unsigned t, unsigned g) {
			constexpr int LEAN = decltype(lc)::value;
			if (o32) {
				if constexpr (two) {
					if (cfg->pattern == TNL_LBM_PATTERN_AA)
						hipLaunchKernelGGL((k_nse_ade<ST, TNL_LBM_COLL_CUM, TNL_LBM_PATTERN_AA, true, LEAN>), dim3(g), dim3(t), 0, st, C, B);
					else
						hipLaunchKernelGGL((k_nse_ade<ST, TNL_LBM_COLL_CUM, TNL_LBM_PATTERN_AB, true, LEAN>), dim3(g), dim3(t), 0, st, C, B);
				}
			}
			else if (cfg->pattern == TNL_LBM_PATTERN_AA)
				hipLaunchKernelGGL((k_nse_ade<ST, TNL_LBM_COLL_CUM, TNL_LBM_PATTERN_AA, false, LEAN>), dim3(g), dim3(t), 0, st, C, B);
			else
				hipLaunchKernelGGL((k_nse_ade<ST, TNL_LBM_COLL_CUM, TNL_LBM_PATTERN_AB, false, LEAN>), dim3(g), dim3(t), 0, st, C, B);
			return hip_check(hipGetLastError(), "nse_ade launch");
		};
		if (A.lists)
			return for_row_lists(A, [&](const StepArgs& C, unsigned t, unsigned g, int lean) {
				if (lean == 1)
					return one(std::integral_constant<int, 1>{}, C, t, g);
				if (lean == 2)
					return one(std::integral_constant<int, 2>{}, C, t, g);
				return one(std::integral_constant<int, 3>{}, C, t, g);
			});
		StepArgs C = A;
		unsigned t = threads, g = grid;
		linearize(C, t, g);
		return one(std::integral_constant<int, 0>{}, C, t, g);
	};
	STORAGE_SWITCH(cfg->storage, go(std::integral_constant<int, ST>{}));
}

int tnl_lbm_d3q7_set_equilibrium(const tnl_lbm_config* cfg, const tnl_lbm_d3q7_data* ADE, const void* phi, double phi0,
								 void* stream)
{
	if (int rc = validate_ade(cfg, ADE))
		return rc;
	const int64_t n = (ADE->X + 2 * ADE->ox) * (ADE->Y + 2 * ADE->oy) * (ADE->Z + 2 * ADE->oz);
	const unsigned grid = (unsigned) ((n + 255) / 256);
	const int ncop = cfg->pattern == TNL_LBM_PATTERN_AB ? 2 : 1;
	hipStream_t st = (hipStream_t) stream;
	auto go = [&](auto stc) {
		constexpr int ST = decltype(stc)::value;
		hipLaunchKernelGGL((k_ade_set_equilibrium<ST>), dim3(grid), dim3(256), 0, st, ADE->df[0], ADE->df[1], ncop, ADE->XYZ, n, phi,
						   cfg->macro_real == TNL_LBM_REAL_F64, phi0, ADE->phi_ref, ADE->store_scale);
		return hip_check(hipGetLastError(), "d3q7 set_equilibrium launch");
	};
	STORAGE_SWITCH(cfg->storage, go(std::integral_constant<int, ST>{}));
}

int tnl_lbm_d3q7_export_f64(const tnl_lbm_config* cfg, const tnl_lbm_d3q7_data* ADE, int which, double* out, void* stream)
{
	if (int rc = validate_ade(cfg, ADE))
		return rc;
	if (which < 0 || which > 1 || ! ADE->df[which] || ! out)
		return fail(TNL_LBM_ERR_ARG, "bad array selector or buffer");
	const int64_t n = (ADE->X + 2 * ADE->ox) * (ADE->Y + 2 * ADE->oy) * (ADE->Z + 2 * ADE->oz);
	const unsigned grid = (unsigned) ((n + 255) / 256);
	hipStream_t st = (hipStream_t) stream;
	auto go = [&](auto stc) {
		constexpr int ST = decltype(stc)::value;
		hipLaunchKernelGGL((k_ade_export<ST>), dim3(grid), dim3(256), 0, st, ADE->df[which], out, ADE->XYZ, n, ADE->phi_ref,
						   ADE->store_scale);
		return hip_check(hipGetLastError(), "d3q7 export launch");
	};
	STORAGE_SWITCH(cfg->storage, go(std::integral_constant<int, ST>{}));
}

static int set_eq_impl(const tnl_lbm_config* cfg, const tnl_lbm_d3q27_data* SD, const double* rho, const double* vx,
					   const double* vy, const double* vz, const double c[4], void* stream)
{
	if (int rc = validate(cfg, SD))
		return rc;
	FieldArgs F;
	F.df[0] = SD->df[0];
	F.df[1] = SD->df[1];
	F.ncopies = cfg->pattern == TNL_LBM_PATTERN_AB ? 2 : 1;
	F.XYZ = SD->XYZ;
	F.nsites = stored_sites(SD);
	F.rho = rho; F.vx = vx; F.vy = vy; F.vz = vz;
	for (int i = 0; i < 4; i++)
		F.c[i] = c[i];
	set_range(SD, F.rs, F.invr);
	const unsigned grid = (unsigned) ((F.nsites + 255) / 256);
	hipStream_t st = (hipStream_t) stream;
	const int eqk = eq_kind_of(cfg->collision);
	auto go = [&](auto stc) {
		constexpr int ST = decltype(stc)::value;
		if (eqk == 0)
			hipLaunchKernelGGL((k_set_equilibrium<ST, 0>), dim3(grid), dim3(256), 0, st, F);
		else if constexpr (! Storage<ST>::dev) {
			if (eqk == 1)
				hipLaunchKernelGGL((k_set_equilibrium<ST, 1>), dim3(grid), dim3(256), 0, st, F);
			else
				hipLaunchKernelGGL((k_set_equilibrium<ST, 2>), dim3(grid), dim3(256), 0, st, F);
		}
		return hip_check(hipGetLastError(), "set_equilibrium launch");
	};
	STORAGE_SWITCH(cfg->storage, go(std::integral_constant<int, ST>{}));
}

int tnl_lbm_d3q27_set_equilibrium(const tnl_lbm_config* cfg, const tnl_lbm_d3q27_data* SD, double rho, double vx,
								  double vy, double vz, void* stream)
{
	const double c[4] = {rho, vx, vy, vz};
	return set_eq_impl(cfg, SD, nullptr, nullptr, nullptr, nullptr, c, stream);
}

int tnl_lbm_d3q27_set_equilibrium_field(const tnl_lbm_config* cfg, const tnl_lbm_d3q27_data* SD, const double* rho,
										const double* vx, const double* vy, const double* vz, void* stream)
{
	if (! rho || ! vx || ! vy || ! vz)
		return fail(TNL_LBM_ERR_NULL, "field pointer missing");
	const double c[4] = {0, 0, 0, 0};
	return set_eq_impl(cfg, SD, rho, vx, vy, vz, c, stream);
}

int tnl_lbm_d3q27_compute_macro(const tnl_lbm_config* cfg, const tnl_lbm_d3q27_data* SD, void* stream)
{
	if (int rc = validate(cfg, SD))
		return rc;
	if (! SD->macro)
		return fail(TNL_LBM_ERR_NULL, "macro array missing");
	const MacroArgs M = macro_args(cfg, SD, SD->df[0]);
	const unsigned grid = (unsigned) ((SD->X * SD->Y * SD->Z + 255) / 256);
	hipStream_t st = (hipStream_t) stream;
	auto go = [&](auto stc) {
		constexpr int ST = decltype(stc)::value;
		hipLaunchKernelGGL((k_compute_macro<ST>), dim3(grid), dim3(256), 0, st, M);
		return hip_check(hipGetLastError(), "compute_macro launch");
	};
	STORAGE_SWITCH(cfg->storage, go(std::integral_constant<int, ST>{}));
}

static int transfer(const tnl_lbm_config* cfg, const tnl_lbm_d3q27_data* SD, int which, double* buf, bool exp, void* stream)
{
	if (int rc = validate(cfg, SD))
		return rc;
	if (which < 0 || which > 1 || ! SD->df[which] || ! buf)
		return fail(TNL_LBM_ERR_ARG, "bad array selector or buffer");
	const int64_t n = stored_sites(SD);
	const unsigned grid = (unsigned) ((n + 255) / 256);
	hipStream_t st = (hipStream_t) stream;
	auto go = [&](auto stc) {
		constexpr int ST = decltype(stc)::value;
		if (exp)
			hipLaunchKernelGGL((k_transfer<ST, 27, true>), dim3(grid), dim3(256), 0, st, SD->df[which], buf, SD->XYZ, n, range_of(SD));
		else
			hipLaunchKernelGGL((k_transfer<ST, 27, false>), dim3(grid), dim3(256), 0, st, SD->df[which], buf, SD->XYZ, n, range_of(SD));
		return hip_check(hipGetLastError(), "transfer launch");
	};
	STORAGE_SWITCH(cfg->storage, go(std::integral_constant<int, ST>{}));
}

int tnl_lbm_d3q27_export_f64(const tnl_lbm_config* cfg, const tnl_lbm_d3q27_data* SD, int which, double* out, void* stream)
{
	return transfer(cfg, SD, which, out, true, stream);
}

int tnl_lbm_d3q27_import_f64(const tnl_lbm_config* cfg, const tnl_lbm_d3q27_data* SD, int which, const double* in, void* stream)
{
	return transfer(cfg, SD, which, const_cast<double*>(in), false, stream);
}

int tnl_lbm_d3q27_halo_x(const tnl_lbm_config* cfg, const tnl_lbm_d3q27_data* lower, const tnl_lbm_d3q27_data* upper,
						 int phase, void* stream)
{
	if (int rc = validate(cfg, lower))
		return rc;
	if (int rc = validate(cfg, upper))
		return rc;
	if (lower->ox < 1 || upper->ox < 1)
		return fail(TNL_LBM_ERR_DIMS, "an x-slab halo needs a ghost plane on each side (ox >= 1)");
	if (lower->Y != upper->Y || lower->Z != upper->Z || lower->oy != upper->oy || lower->oz != upper->oz)
		return fail(TNL_LBM_ERR_DIMS, "neighbouring slabs must have the same Y, Z, oy and oz");
	if (phase != TNL_LBM_HALO_PULL && phase != TNL_LBM_HALO_PUSH)
		return fail(TNL_LBM_ERR_ARG, "unknown halo phase");
	const bool aa = cfg->pattern == TNL_LBM_PATTERN_AA;
	if (phase == TNL_LBM_HALO_PUSH && ! aa)
		return fail(TNL_LBM_ERR_UNSUPPORTED, "the push phase belongs to the A-A odd step");
	const uint64_t eb = (uint64_t) ELEM_BYTES[cfg->storage];
	const int64_t plane = (lower->Y + 2 * lower->oy) * (lower->Z + 2 * lower->oz);
	auto at = [&](const tnl_lbm_d3q27_data* D, int q, int64_t xs) {
		return (char*) D->df[0] + ((uint64_t) q * (uint64_t) D->XYZ + (uint64_t) (xs * plane)) * eb;
	};
	// storage x of the planes next to the face
	const int64_t lo_last = lower->ox + lower->X - 1, lo_ghost = lower->ox + lower->X;
	const int64_t up_first = upper->ox, up_ghost = upper->ox - 1;
	hipStream_t st = (hipStream_t) stream;
	for (int q = 0; q < 27; q++) {
		if (CX[q] == 0)
			continue;
		// Lower's ghost plane (its +x side) takes from upper's first plane the
		// slots with c_x = +1 for A-A (the odd step reads slot opp(q) of the
		// populations q that move in -x) and with c_x = -1 for A-B (pulled as
		// they are); the other nine slots go from lower's last plane to
		// upper's ghost plane.  The push phase runs the same pairs backwards.
		const bool up_to_lo = aa ? CX[q] > 0 : CX[q] < 0;
		char *dst, *src;
		if (phase == TNL_LBM_HALO_PULL) {
			dst = up_to_lo ? at(lower, q, lo_ghost) : at(upper, q, up_ghost);
			src = up_to_lo ? at(upper, q, up_first) : at(lower, q, lo_last);
		}
		else {
			dst = up_to_lo ? at(upper, q, up_first) : at(lower, q, lo_last);
			src = up_to_lo ? at(lower, q, lo_ghost) : at(upper, q, up_ghost);
		}
		if (int rc = hip_check(hipMemcpyAsync(dst, src, (size_t) plane * eb, hipMemcpyDeviceToDevice, st), "halo copy"))
			return rc;
	}
	return TNL_LBM_OK;
}

// Partial sums of the FP64 reductions: one buffer per (host thread,
// device), allocated on that thread's first reduction on the device and held
// until the process ends (RED_BLOCKS + 1 rows of RED_N doubles, 49 KB).
// Each reduction synchronises its stream before it returns, so the next one
// of the same thread may reuse the buffer on any stream.
static int partial_buffer(double** out)
{
	constexpr int MAX_DEVICES = 64;
	static thread_local double* bufs[MAX_DEVICES] = {};
	int dev = 0;
	if (int rc = hip_check(hipGetDevice(&dev), "hipGetDevice"))
		return rc;
	if (dev < 0 || dev >= MAX_DEVICES)
		return fail(TNL_LBM_ERR_ARG, "device ordinal beyond the reduction buffers' table");
	if (! bufs[dev])
		if (int rc = hip_check(hipMalloc(&bufs[dev], sizeof(double) * (RED_BLOCKS * RED_N + RED_N)), "hipMalloc reduce buffer"))
			return rc;
	*out = bufs[dev];
	return TNL_LBM_OK;
}

// partial sums -> out (device) on `st`
static int finish_reduction(double* partial, double* out_dev, hipStream_t st)
{
	hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(256), 0, st, (const double*) partial, RED_BLOCKS, out_dev);
	return hip_check(hipGetLastError(), "reduce launch");
}

static int reduction_args(const tnl_lbm_config* cfg, const tnl_lbm_d3q27_data* SD, MacroArgs& M)
{
	if (int rc = validate(cfg, SD))
		return rc;
	if (! SD->macro)
		return fail(TNL_LBM_ERR_NULL, "macro array missing");
	M = macro_args(cfg, SD, SD->df[0]);
	M.dist_x = M.dist_y = M.dist_z = 0;
	return TNL_LBM_OK;
}

int tnl_lbm_d3q27_reduce_macro(const tnl_lbm_config* cfg, const tnl_lbm_d3q27_data* SD, double out[6], void* stream)
{
	MacroArgs M;
	if (int rc = reduction_args(cfg, SD, M))
		return rc;
	if (! out)
		return fail(TNL_LBM_ERR_NULL, "output missing");
	double* partial = nullptr;
	if (int rc = partial_buffer(&partial))
		return rc;
	hipStream_t st = (hipStream_t) stream;
	hipLaunchKernelGGL(k_reduce_macro, dim3(RED_BLOCKS), dim3(256), 0, st, M, partial);
	if (int rc = finish_reduction(partial, partial + RED_BLOCKS * RED_N, st))
		return rc;
	if (int rc = hip_check(hipMemcpyAsync(out, partial + RED_BLOCKS * RED_N, sizeof(double) * RED_N, hipMemcpyDeviceToHost, st), "reduce copy"))
		return rc;
	return hip_check(hipStreamSynchronize(st), "reduce sync");
}

int tnl_lbm_d3q27_reduce_macro_noalloc(const tnl_lbm_config* cfg, const tnl_lbm_d3q27_data* SD, double* out_device,
									   void* stream)
{
	MacroArgs M;
	if (int rc = reduction_args(cfg, SD, M))
		return rc;
	if (! out_device)
		return fail(TNL_LBM_ERR_NULL, "device output missing");
	// every one of the RED_BLOCKS workgroups writes its own partial row, so
	// nothing in the scratch needs clearing; the grid never exceeds the rows
	static_assert(RED_BLOCKS <= SCRATCH_ROWS, "the reduction grid must fit the scratch");
	const int blocks = RED_BLOCKS;
	if (blocks > SCRATCH_ROWS)
		return fail(TNL_LBM_ERR_DIMS, "reduction grid larger than the scratch (1024 partial rows)");
	double* partial = nullptr;
	if (int rc = scratch_address(&partial))
		return rc;
	hipStream_t st = (hipStream_t) stream;
	hipLaunchKernelGGL(k_reduce_macro, dim3(blocks), dim3(256), 0, st, M, partial);
	return finish_reduction(partial, out_device, st);
}

int tnl_lbm_d3q27_reduce_gradients(const tnl_lbm_config* cfg, const tnl_lbm_d3q27_data* SD, uint32_t flags, double out[2],
								   void* stream)
{
	MacroArgs M;
	if (int rc = reduction_args(cfg, SD, M))
		return rc;
	if (! out)
		return fail(TNL_LBM_ERR_NULL, "output missing");
	M.dist_x = (flags & TNL_LBM_FLAG_DISTRIBUTED_X) != 0;
	M.dist_y = (flags & TNL_LBM_FLAG_DISTRIBUTED_Y) != 0;
	M.dist_z = (flags & TNL_LBM_FLAG_DISTRIBUTED_Z) != 0;
	if ((M.dist_x && SD->ox < 1) || (M.dist_y && SD->oy < 1) || (M.dist_z && SD->oz < 1))
		return fail(TNL_LBM_ERR_DIMS, "a distributed axis needs an overlap of at least one site");
	double* partial = nullptr;
	if (int rc = partial_buffer(&partial))
		return rc;
	hipStream_t st = (hipStream_t) stream;
	hipLaunchKernelGGL(k_reduce_gradients, dim3(RED_BLOCKS), dim3(256), 0, st, M, partial);
	if (int rc = finish_reduction(partial, partial + RED_BLOCKS * RED_N, st))
		return rc;
	double res[RED_N];
	if (int rc = hip_check(hipMemcpyAsync(res, partial + RED_BLOCKS * RED_N, sizeof(double) * RED_N, hipMemcpyDeviceToHost, st), "reduce copy"))
		return rc;
	if (int rc = hip_check(hipStreamSynchronize(st), "reduce sync"))
		return rc;
	out[0] = res[0];
	out[1] = res[1];
	return TNL_LBM_OK;
}

}  // extern "C"

// the configurations the adjoint and the moment recording are built for
static int adjoint_config(const tnl_lbm_config* cfg, const tnl_lbm_d3q27_data* SD)
{
	if (! cfg || ! SD)
		return fail(TNL_LBM_ERR_NULL, "null config or data");
	if (cfg->collision != TNL_LBM_COLL_CUM || cfg->pattern != TNL_LBM_PATTERN_AB)
		return fail(TNL_LBM_ERR_UNSUPPORTED, "the adjoint is built for the default cumulant with the A-B pattern");
	if (cfg->storage != TNL_LBM_STORE_F64 && cfg->storage != TNL_LBM_STORE_F32)
		return fail(TNL_LBM_ERR_UNSUPPORTED, "the adjoint is built for the absolute storage forms (F64, F32)");
	if (SD->ox || SD->oy || SD->oz)
		return fail(TNL_LBM_ERR_UNSUPPORTED, "the adjoint is built for a single block without overlaps");
	if (SD->force_field)
		return fail(TNL_LBM_ERR_UNSUPPORTED, "the adjoint is built for the homogeneous force only (no force_field)");
	return TNL_LBM_OK;
}

static int tape_kind_of(const tnl_lbm_config* cfg, const void* tape, int tape_f64)
{
	if (! tape)
		return fail(TNL_LBM_ERR_NULL, "moment tape missing");
	if (tape_f64 && cfg->storage != TNL_LBM_STORE_F64)
		return fail(TNL_LBM_ERR_UNSUPPORTED, "a double tape needs F64 storage");
	return tape_f64 ? 2 : 1;
}

static int adjoint_impl(const tnl_lbm_config* cfg, const tnl_lbm_d3q27_data* SD, const void* tape, int tape_kind,
						const void* lambda_next, void* lambda, double* grad_inflow, void* stream)
{
	if (! lambda_next || ! lambda)
		return fail(TNL_LBM_ERR_NULL, "adjoint arrays missing");
	tnl_lbm_d3q27_data D = *SD;
	if (tape_kind) {
		// the state is not read: any non-null placeholder passes the layout checks
		D.df[0] = D.df[1] = (void*) tape;
		if (lambda == lambda_next || lambda == tape)
			return fail(TNL_LBM_ERR_ARG, "lambda must not alias lambda_next or the tape");
	}
	else {
		if (lambda == lambda_next || lambda == SD->df[0])
			return fail(TNL_LBM_ERR_ARG, "lambda must not alias lambda_next or the state");
		if (! D.df[1])
			D.df[1] = D.df[0];  // only df[0] is read
	}
	StepArgs A;
	unsigned threads = 0, grid = 0;
	const int64_t off[3] = {0, 0, 0}, end[3] = {SD->X, SD->Y, SD->Z};
	if (int rc = make_step_args(cfg, &D, off, end, 0, A, threads, grid))
		return rc;
	if (grid == 0)
		return TNL_LBM_OK;
	const AdjPtrs P{lambda_next, lambda, grad_inflow, tape};
	return launch_adjoint(cfg, A, P, tape_kind, threads, grid, (hipStream_t) stream);
}

int tnl_lbm_d3q27_adjoint_step(const tnl_lbm_config* cfg, const tnl_lbm_d3q27_data* SD, const void* lambda_next,
							   void* lambda, double* grad_inflow, void* stream)
{
	if (int rc = adjoint_config(cfg, SD))
		return rc;
	return adjoint_impl(cfg, SD, nullptr, 0, lambda_next, lambda, grad_inflow, stream);
}

int tnl_lbm_d3q27_step_record(const tnl_lbm_config* cfg, const tnl_lbm_d3q27_data* SD, void* tape, int tape_f64,
							  void* stream)
{
	if (int rc = adjoint_config(cfg, SD))
		return rc;
	const int kind = tape_kind_of(cfg, tape, tape_f64);
	if (kind < 0)
		return kind;
	if (tape == SD->df[0] || tape == SD->df[1])
		return fail(TNL_LBM_ERR_ARG, "the tape must not alias the populations");
	StepArgs A;
	unsigned threads = 0, grid = 0;
	const int64_t off[3] = {0, 0, 0}, end[3] = {SD->X, SD->Y, SD->Z};
	if (int rc = make_step_args(cfg, SD, off, end, 0, A, threads, grid))
		return rc;
	if (grid == 0)
		return TNL_LBM_OK;
	A.tape = tape;
	return launch_record(cfg, A, kind, threads, grid, (hipStream_t) stream);
}

int tnl_lbm_d3q27_adjoint_step_recorded(const tnl_lbm_config* cfg, const tnl_lbm_d3q27_data* SD, const void* tape,
										int tape_f64, const void* lambda_next, void* lambda, double* grad_inflow,
										void* stream)
{
	if (int rc = adjoint_config(cfg, SD))
		return rc;
	const int kind = tape_kind_of(cfg, tape, tape_f64);
	if (kind < 0)
		return kind;
	return adjoint_impl(cfg, SD, tape, kind, lambda_next, lambda, grad_inflow, stream);
}

int tnl_lbm_d3q7_adjoint_step(const tnl_lbm_config* cfg, const tnl_lbm_d3q27_data* nse, const tnl_lbm_d3q7_data* ade,
							  const void* velocity, const void* lambda_next, void* lambda, double* grad_source,
							  void* stream)
{
	if (! cfg || ! nse || ! ade)
		return fail(TNL_LBM_ERR_NULL, "null config or data");
	if (cfg->pattern != TNL_LBM_PATTERN_AB)
		return fail(TNL_LBM_ERR_UNSUPPORTED, "the D3Q7 adjoint is built for the A-B pattern");
	if (nse->ox || nse->oy || nse->oz)
		return fail(TNL_LBM_ERR_UNSUPPORTED, "the D3Q7 adjoint is built for a single block without overlaps");
	if (ade->struct_size != sizeof(tnl_lbm_d3q7_data) || ade->version != TNL_LBM_ABI_VERSION)
		return fail(TNL_LBM_ERR_VERSION, "tnl_lbm_d3q7_data struct_size/version mismatch");
	if (ade->collision < TNL_LBM_ADE_COLL_SRT || ade->collision > TNL_LBM_ADE_COLL_CLBM)
		return fail(TNL_LBM_ERR_UNSUPPORTED, "unknown ADE collision");
	if (ade->X != nse->X || ade->Y != nse->Y || ade->Z != nse->Z || ade->ox || ade->oy || ade->oz || ade->XYZ != nse->XYZ)
		return fail(TNL_LBM_ERR_DIMS, "NSE and ADE lattices must have the same geometry");
	if (! velocity || ! lambda_next || ! lambda)
		return fail(TNL_LBM_ERR_NULL, "velocity or adjoint arrays missing");
	if (lambda == lambda_next)
		return fail(TNL_LBM_ERR_ARG, "lambda must not alias lambda_next");
	tnl_lbm_d3q27_data D = *nse;
	D.df[0] = D.df[1] = (void*) velocity;  // the populations are not read: placeholders for the layout checks
	StepArgs A;
	unsigned threads = 0, grid = 0;
	const int64_t off[3] = {0, 0, 0}, end[3] = {nse->X, nse->Y, nse->Z};
	if (int rc = make_step_args(cfg, &D, off, end, 0, A, threads, grid))
		return rc;
	if (grid == 0)
		return TNL_LBM_OK;
	const AdeAdjArgs P{velocity, lambda_next, lambda, grad_source, ade->map, ade->diffusion_coefficient,
					   ade->lbm_diffusivity, ade->collision};
	return launch_ade_adjoint(cfg, A, P, threads, grid, (hipStream_t) stream);
}
