/*
 * tnl_lbm_amd.h -- C ABI of the MI355X (gfx950) lattice-update library.
 *
 * One shared library, libtnl_lbm_amd.so, replaces the reference's fused
 * per-cell kernels (TNL-LBM include/lbm3d/kernels.h:59-196) and the device
 * loops around them (include/lbm3d/lbm_block.hpp:248-305,
 * include/lbm3d/state.hpp:1235-1409).  The reference instantiates one
 * template per LBM_CONFIG<...> bundle (include/lbm3d/defs.h:288-310); a C ABI
 * cannot carry templates, so every entry point here takes a small
 * tnl_lbm_config that names the plugins at run time and the library
 * dispatches to a kernel compiled for that combination.
 *
 * Conventions
 *  - Plain C types only; all array pointers are DEVICE pointers owned by the
 *    caller.  The library never allocates on a step path.
 *  - Every function returns 0 (TNL_LBM_OK) or a negative status code and
 *    never throws; tnl_lbm_last_error() returns a thread-local message.
 *  - `stream` is a hipStream_t passed as void* (NULL = the null stream).
 *    Calls are asynchronous on that stream.
 *  - Lattice layout is the reference's SoA layout (defs.h:90,
 *    lbm_data.h:50-54): population q of site (x,y,z) lives at
 *        q*XYZ + ((x+ox)*(Z+2oz) + (z+oz))*(Y+2oy) + (y+oy)
 *    i.e. y is contiguous, then z, then x, then q.  With zero overlaps (one
 *    GPU) this is q*XYZ + (x*Z + z)*Y + y.  The macro array uses the same
 *    layout with the quantity id in place of q (lbm_data.h:56-60).
 *  - Direction numbering is the reference enum (defs.h:317-350).
 */
#ifndef TNL_LBM_AMD_H
#define TNL_LBM_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TNL_LBM_ABI_VERSION 4u

/* ---- status codes ------------------------------------------------------ */
#define TNL_LBM_OK 0
#define TNL_LBM_ERR_NULL (-1)        /* a required pointer is NULL            */
#define TNL_LBM_ERR_DIMS (-2)        /* sizes / ranges are inconsistent        */
#define TNL_LBM_ERR_UNSUPPORTED (-3) /* plugin combination not built          */
#define TNL_LBM_ERR_VERSION (-4)     /* struct_size / version mismatch         */
#define TNL_LBM_ERR_HIP (-5)         /* a HIP runtime call failed              */
#define TNL_LBM_ERR_ARG (-6)         /* other invalid argument                 */

/* ---- population directions (defs.h:317-350) --------------------------- */
enum tnl_lbm_direction {
	TNL_LBM_zzz = 0, TNL_LBM_pzz = 1, TNL_LBM_mzz = 2, TNL_LBM_zpz = 3, TNL_LBM_zmz = 4,
	TNL_LBM_zzp = 5, TNL_LBM_zzm = 6, TNL_LBM_ppz = 7, TNL_LBM_mmz = 8, TNL_LBM_pmz = 9,
	TNL_LBM_mpz = 10, TNL_LBM_pzp = 11, TNL_LBM_mzm = 12, TNL_LBM_pzm = 13, TNL_LBM_mzp = 14,
	TNL_LBM_zpp = 15, TNL_LBM_zmm = 16, TNL_LBM_zpm = 17, TNL_LBM_zmp = 18, TNL_LBM_ppp = 19,
	TNL_LBM_mmm = 20, TNL_LBM_ppm = 21, TNL_LBM_mmp = 22, TNL_LBM_pmp = 23, TNL_LBM_mpm = 24,
	TNL_LBM_pmm = 25, TNL_LBM_mpp = 26
};

/* ---- D3Q27 geometry codes: D3Q27_BC_All::GEO (d3q27/bc.h:17-43) -------- */
enum tnl_lbm_d3q27_geo {
	TNL_LBM_GEO_FLUID = 0,
	TNL_LBM_GEO_WALL = 1,
	TNL_LBM_GEO_INFLOW = 2,
	TNL_LBM_GEO_INFLOW_LEFT = 3,
	TNL_LBM_GEO_INFLOW_BB_LEFT = 4,
	TNL_LBM_GEO_INFLOW_EQ_LEFT = 5,
	TNL_LBM_GEO_OUTFLOW_EQ = 6,
	TNL_LBM_GEO_OUTFLOW_RIGHT = 7,
	TNL_LBM_GEO_OUTFLOW_RIGHT_INTERP = 8,
	TNL_LBM_GEO_PERIODIC = 9,
	TNL_LBM_GEO_NOTHING = 10,
	TNL_LBM_GEO_SYM_TOP = 11,
	TNL_LBM_GEO_SYM_BOTTOM = 12,
	TNL_LBM_GEO_SYM_LEFT = 13,
	TNL_LBM_GEO_SYM_RIGHT = 14,
	TNL_LBM_GEO_SYM_BACK = 15,
	TNL_LBM_GEO_SYM_FRONT = 16
};

/* ---- D3Q7 geometry codes: D3Q7_BC_All::GEO (d3q7/bc.h:17-36) ----------- */
enum tnl_lbm_d3q7_geo {
	TNL_LBM_ADE_GEO_FLUID = 0,
	TNL_LBM_ADE_GEO_WALL = 1,
	TNL_LBM_ADE_GEO_WALL_BODY = 2,
	TNL_LBM_ADE_GEO_SOLID = 3,
	TNL_LBM_ADE_GEO_INFLOW = 7,
	TNL_LBM_ADE_GEO_OUTFLOW_RIGHT = 8,
	TNL_LBM_ADE_GEO_PERIODIC = 9,
	TNL_LBM_ADE_GEO_NOTHING = 10,
	TNL_LBM_ADE_GEO_SYM_TOP = 12,
	TNL_LBM_ADE_GEO_SYM_BOTTOM = 13,
	TNL_LBM_ADE_GEO_SYM_LEFT = 14,
	TNL_LBM_ADE_GEO_SYM_RIGHT = 15,
	TNL_LBM_ADE_GEO_SYM_BACK = 16,
	TNL_LBM_ADE_GEO_SYM_FRONT = 17
};

/* ---- storage forms of the population array ----------------------------
 * F64      : absolute f in double, FP64 arithmetic (TraitsDP, defs.h:177)
 * F32      : absolute f in float,  FP32 arithmetic (TraitsSP, defs.h:176)
 * F32_DEV  : deviation f - w_q stored in float, FP32 arithmetic
 * F16_DEV  : deviation f - w_q stored in IEEE binary16, FP32 arithmetic
 * BF16_DEV : deviation f - w_q stored in bfloat16, FP32 arithmetic
 * I16_DEV  : deviation as 16-bit fixed point: the signed-normalised integer
 *            round(32767 * clamp((f - w_q) / (w_q * R), -1, 1)) with R =
 *            store_range of the data struct; FP32 arithmetic.  Same 2 bytes
 *            per population as FP16 with a uniform resolution of w_q R/32767.
 * The deviation forms follow the "well-conditioned" family of the reference
 * (d3q27/common_well.h:37, d3q27/eq_well.h:21-134), which is not wired into
 * any reference configuration; here it is a first-class storage form.      */
enum tnl_lbm_storage {
	TNL_LBM_STORE_F64 = 0,
	TNL_LBM_STORE_F32 = 1,
	TNL_LBM_STORE_F32_DEV = 2,
	TNL_LBM_STORE_F16_DEV = 3,
	TNL_LBM_STORE_BF16_DEV = 4,
	TNL_LBM_STORE_I16_DEV = 5
};

/* ---- macro / field element types -------------------------------------- */
enum tnl_lbm_real { TNL_LBM_REAL_F64 = 0, TNL_LBM_REAL_F32 = 1 };

/* ---- streaming patterns ------------------------------------------------
 * AA : in-place A-A even/odd pattern (d3q27/streaming_AA.h:7-118), one copy
 * AB : pull scheme with two copies   (d3q27/streaming_AB.h:7-60)           */
enum tnl_lbm_pattern { TNL_LBM_PATTERN_AA = 0, TNL_LBM_PATTERN_AB = 1 };

/* ---- D3Q27 collision operators (the COLL plugin) ----------------------- */
enum tnl_lbm_collision {
	TNL_LBM_COLL_CUM = 0,               /* D3Q27_CUM, default flags     col_cum.h:13-511        */
	TNL_LBM_COLL_CUM_2017 = 1,          /* + USE_GEIER_CUM_2017          col_cum.h:112-329       */
	TNL_LBM_COLL_CUM_2017_ANTIALIAS = 2,/* + USE_GEIER_CUM_ANTIALIAS     col_cum.h:234-245       */
	TNL_LBM_COLL_MRT_LES = 3,           /* D3Q27_MRT ("MRT_LES")         col_mrt.h:7-142         */
	TNL_LBM_COLL_KBC_N1 = 4,            /* D3Q27_KBC_N1..N4              col_kbc_n.h:254-1272    */
	TNL_LBM_COLL_KBC_N2 = 5,
	TNL_LBM_COLL_KBC_N3 = 6,
	TNL_LBM_COLL_KBC_N4 = 7,
	TNL_LBM_COLL_KBC_C1 = 8,            /* D3Q27_KBC_C1..C4              col_kbc_c.h:283-1301    */
	TNL_LBM_COLL_KBC_C2 = 9,
	TNL_LBM_COLL_KBC_C3 = 10,
	TNL_LBM_COLL_KBC_C4 = 11
};

/* ---- D3Q7 collision operators ------------------------------------------ */
enum tnl_lbm_ade_collision {
	TNL_LBM_ADE_COLL_SRT = 0,  /* D3Q7_SRT  d3q7/col_srt.h:8-27   */
	TNL_LBM_ADE_COLL_MRT = 1,  /* D3Q7_MRT  d3q7/col_mrt.h:8-51   */
	TNL_LBM_ADE_COLL_CLBM = 2  /* D3Q7_CLBM d3q7/col_clbm.h:10-85 */
};

/* ---- step flags --------------------------------------------------------- */
#define TNL_LBM_FLAG_OUTPUT_MACRO 0x1u  /* output_macro argument (kernels.h:66,110) */
#define TNL_LBM_FLAG_DISTRIBUTED_X 0x2u /* bool3d distributed (kernels.h:9,22-27)  */
#define TNL_LBM_FLAG_DISTRIBUTED_Y 0x4u
#define TNL_LBM_FLAG_DISTRIBUTED_Z 0x8u

/* Plugin bundle: the run-time image of LBM_CONFIG<TRAITS, KernelStruct,
 * DATA, COLL, EQ, STREAMING, BC, MACRO> (defs.h:288-310).  EQ is the one
 * the collision implies (D3Q27_EQ_INV_CUM for the cumulants, D3Q27_EQ for
 * MRT, the entropic equilibrium for KBC); BC is D3Q27_BC_All; MACRO is
 * D3Q27_MACRO_Default (rho, vx, vy, vz).                                   */
typedef struct tnl_lbm_config {
	uint32_t struct_size; /* sizeof(tnl_lbm_config) */
	uint32_t version;     /* TNL_LBM_ABI_VERSION    */
	int32_t collision;    /* enum tnl_lbm_collision */
	int32_t storage;      /* enum tnl_lbm_storage   */
	int32_t pattern;      /* enum tnl_lbm_pattern   */
	int32_t macro_real;   /* enum tnl_lbm_real: element type of `macro` */
} tnl_lbm_config;

/* Kernel data: the run-time image of NSE_Data_ConstInflow /
 * NSE_Data_InflowProfile (lbm_data.h:8-133), passed by pointer.            */
typedef struct tnl_lbm_d3q27_data {
	uint32_t struct_size; /* sizeof(tnl_lbm_d3q27_data) */
	uint32_t version;     /* TNL_LBM_ABI_VERSION        */
	int64_t X, Y, Z;      /* local sizes without overlaps (LBM_Data::X/Y/Z) */
	int64_t ox, oy, oz;   /* overlaps (indexer overlaps; 0 on one GPU)      */
	int64_t XYZ;          /* population stride; >= (X+2ox)(Y+2oy)(Z+2oz)    */
	int32_t even_iter;    /* A-A parity, LBM::updateKernelData (lbm.hpp:311) */
	int32_t inflow_profile_real; /* enum tnl_lbm_real: element type of inflow_profile
	                                (F64 = 0, the default; F32 for TraitsSP's dreal) */
	double lbm_viscosity; /* LBM_Data::lbmViscosity                          */
	double fx, fy, fz;    /* NSE_Data homogeneous force (lbm_data.h:69-72)   */
	void* df[2];          /* dfs[df_cur], dfs[df_out] (df_out: A-B only)     */
	void* macro;          /* dmacro: 4*XYZ elements of macro_real, or NULL   */
	const int16_t* map;   /* dmap: XYZ geometry codes; NULL = all PERIODIC   */
	double inflow_vx, inflow_vy, inflow_vz; /* NSE_Data_ConstInflow (lbm_data.h:81-91) */
	/* NSE_Data_InflowProfile (lbm_data.h:100-132): device arrays of
	 * inflow_profile_real elements indexed Y*z + y; a non-NULL pointer
	 * overrides the constant component.                                    */
	const void* inflow_profile[3];
	double store_range;   /* I16_DEV only: R, the bound of |f - w_q| / w_q (> 0) */
	/* I16_DEV only: the bound per direction class |c_q|^2 = 0, 1, 2, 3 (rest,
	 * faces, edges, corners); a class with 0 uses store_range.  The classes'
	 * deviations differ by orders of magnitude (the rest population's is
	 * ~1.5|u|^2), so per-class bounds buy up to 6 bits of resolution.       */
	double store_range_class[4];
	/* MACRO::computeForcing (d3q27/macro.h:50-52, called at kernels.h:103 and
	 * :162): a per-site force added to fx, fy, fz before the collision -- three
	 * planes (fx, fy, fz) of XYZ elements of cfg->macro_real in the lattice
	 * layout, or NULL for none.  This is the form in which the reference's
	 * IBM (the macro quantities e_fx..e_fz, lagrange_3D.hpp) and
	 * non-Newtonian (nonNewtonian.h:679-760) force terms reach the kernel.
	 * The map kernel and the coupled kernel read it; a box without a map and
	 * with a force field is stepped by the map kernel as all PERIODIC.  The
	 * adjoint entry points do not take it (TNL_LBM_ERR_UNSUPPORTED).         */
	const void* force_field;
	/* Row lists of this map (no counterpart in the reference; NULL = none):
	 * the output of tnl_lbm_classify_rows(SD, NULL, ...).  A step over the
	 * whole box (offset 0, end = X, Y, Z) of the default cumulant in a
	 * deviation storage form then runs each list through its own kernel:
	 * rows of FLUID and PERIODIC sites only, rows that also hold WALL sites,
	 * and the remaining rows through the kernel with every boundary code
	 * (the F64 and F32 forms gain nothing from it and ignore the lists).  The
	 * lists must describe the current map: classify again after changing it. */
	const uint32_t* rows;
	int64_t row_count[3];
} tnl_lbm_d3q27_data;

/* D3Q7 ADE data: ADE_Data (lbm_data.h:212-246) plus a source term that the
 * reference leaves as a TODO (lbm_data.h:229).                             */
typedef struct tnl_lbm_d3q7_data {
	uint32_t struct_size; /* sizeof(tnl_lbm_d3q7_data) */
	uint32_t version;
	int64_t X, Y, Z;
	int64_t ox, oy, oz;
	int64_t XYZ;
	int32_t even_iter;
	int32_t collision;       /* enum tnl_lbm_ade_collision                    */
	double lbm_diffusivity;  /* ADE_Data::lbmViscosity (the diffusivity D)    */
	void* df[2];             /* 7*XYZ elements of the NSE storage form        */
	void* macro;             /* phi: XYZ elements of macro_real, or NULL      */
	const int16_t* map;      /* D3Q7 geometry codes; NULL = all PERIODIC       */
	const void* diffusion_coefficient; /* per-site D (macro_real), or NULL    */
	const void* source;      /* per-site source S (macro_real), or NULL; adds
	                            w_q*S to every population after collision     */
	double inflow_phi;       /* ADE_Data_ConstInflow::inflow_phi               */
	/* deviation storage forms only: stored = (f_q - w_q*phi_ref)*store_scale */
	double phi_ref;
	double store_scale;
	/* Row lists of the coupled step (no counterpart in the reference; NULL =
	 * none): the output of tnl_lbm_classify_rows(nse, ade, ...) for this map
	 * and the NSE lattice's, used by tnl_lbm_nse_ade_step over the whole box. */
	const uint32_t* rows;
	int64_t row_count[3];
} tnl_lbm_d3q7_data;

/* ======================================================================= */
/* library information                                                     */
uint32_t tnl_lbm_abi_version(void);
const char* tnl_lbm_last_error(void);

/* No counterpart in the reference: a diagnostic switch.  Where a population
 * array holds at most 4 GiB, the map kernel (2-byte forms), the coupled
 * kernel (2-byte forms), the recording step and the FP64 adjoint address
 * each slot as a per-direction base plus a 32-bit offset; on = 1 makes every
 * launch of this process take the 64-bit form (same results, bit for bit),
 * on = 0 restores the default.  Returns the previous setting.              */
int tnl_lbm_set_wide_addressing(int on);

/* No counterpart in the reference: classification of the interior (x, z)
 * rows for the row-list launches (the rows, row_count fields above).  A row
 * is class 1 when every site of it is FLUID or PERIODIC (and, with ade, every
 * D3Q7 code FLUID or PERIODIC), class 2 when WALL sites occur as well, class
 * 0 otherwise; a NULL map counts as all PERIODIC.  rows: device array of at
 * least X*Z entries (caller-owned), filled with the rows of class 0, then
 * class 1, then class 2, each in increasing x*Z + z order, each entry
 * x << 16 | z; row_count: host array, receives the three counts.  Needs
 * X, Z < 65536.  Allocates temporary device memory and waits for the stream
 * (a set-up call, not a step-path one).                                    */
int tnl_lbm_classify_rows(const tnl_lbm_d3q27_data* nse, const tnl_lbm_d3q7_data* ade, uint32_t* rows,
						  int64_t row_count[3], void* stream);

/* ---- D3Q27 lattice update ----------------------------------------------
 * Replaces one launch of cudaLBMKernel<NSE>(SD, offset, end, distributed,
 * output_macro) (kernels.h:59-112, launched at state.hpp:1307-1309) over the
 * box offset <= (x,y,z) < end.  flags: TNL_LBM_FLAG_*.                     */
int tnl_lbm_d3q27_step(const tnl_lbm_config* cfg, const tnl_lbm_d3q27_data* SD,
                       const int64_t offset[3], const int64_t end[3], uint32_t flags, void* stream);

/* ---- streaming-only measurement step -----------------------------------
 * No counterpart in the reference: a measurement aid.  One A-A step of free
 * streaming (the collision replaced by the identity) over the whole box,
 * run by the same map-free periodic kernels, loads, stores and launch shape
 * as tnl_lbm_d3q27_step for this storage form and SD->even_iter: its time is
 * the memory ceiling of the step's access pattern on this allocation.  Even
 * step: slot opp(q) of each site takes slot q; odd step: slot q at r + c_q
 * takes slot opp(q) at r - c_q (two steps move population q by 2 c_q).
 * Needs pattern A-A, no map and no force field.                            */
int tnl_lbm_d3q27_stream_only(const tnl_lbm_config* cfg, const tnl_lbm_d3q27_data* SD, void* stream);

/* ---- coupled NSE + ADE update ------------------------------------------
 * Replaces cudaLBMKernel<NSE,ADE> (kernels.h:114-196): the D3Q27 step, then
 * the D3Q7 step at the same site driven by the NSE velocity.               */
int tnl_lbm_nse_ade_step(const tnl_lbm_config* cfg, const tnl_lbm_d3q27_data* nse,
                         const tnl_lbm_d3q7_data* ade, const int64_t offset[3], const int64_t end[3],
                         uint32_t flags, void* stream);

/* ---- initialisation -----------------------------------------------------
 * set_equilibrium: LBM_BLOCK::setEquilibrium(rho, vx, vy, vz)
 *   (lbm_block.hpp:248-278): f = f_eq(rho,u) at every stored site, including
 *   overlaps; with A-B both copies are written.
 * set_equilibrium_field: COLL::setEquilibriumLat over a field
 *   (common.h:126-158, used by the initial conditions, sim_4.cu:145-163):
 *   rho/vx/vy/vz are DEVICE double arrays of XYZ elements in the lattice
 *   layout.
 * compute_macro: LBM_BLOCK::computeInitialMacro (lbm_block.hpp:281-305).  */
int tnl_lbm_d3q27_set_equilibrium(const tnl_lbm_config* cfg, const tnl_lbm_d3q27_data* SD,
                                  double rho, double vx, double vy, double vz, void* stream);
int tnl_lbm_d3q27_set_equilibrium_field(const tnl_lbm_config* cfg, const tnl_lbm_d3q27_data* SD,
                                        const double* rho, const double* vx, const double* vy,
                                        const double* vz, void* stream);
int tnl_lbm_d3q27_compute_macro(const tnl_lbm_config* cfg, const tnl_lbm_d3q27_data* SD, void* stream);

/* ---- population transfer ------------------------------------------------
 * Convert between the stored form (storage dtype, deviation offset) and
 * absolute double populations in the same q*XYZ layout, slot by slot (no
 * A-A reordering).  `out`/`in` are device arrays of 27*XYZ doubles;
 * `which` selects df[0] or df[1].                                          */
int tnl_lbm_d3q27_export_f64(const tnl_lbm_config* cfg, const tnl_lbm_d3q27_data* SD, int which,
                             double* out, void* stream);
int tnl_lbm_d3q27_import_f64(const tnl_lbm_config* cfg, const tnl_lbm_d3q27_data* SD, int which,
                             const double* in, void* stream);

/* ---- D3Q7 initialisation (ADE) ------------------------------------------
 * f_q = D3Q7_EQ::eq_q(phi, 0) (d3q7/common.h:105-115) at every stored site:
 * phi is a device array of XYZ elements (macro_real) or NULL for phi0.
 * export_f64 returns absolute populations (7*XYZ doubles).                 */
int tnl_lbm_d3q7_set_equilibrium(const tnl_lbm_config* cfg, const tnl_lbm_d3q7_data* ADE,
                                 const void* phi, double phi0, void* stream);
int tnl_lbm_d3q7_export_f64(const tnl_lbm_config* cfg, const tnl_lbm_d3q7_data* ADE, int which,
                            double* out, void* stream);

/* ---- FP64 global reductions ---------------------------------------------
 * Sums over the local box of macro quantities, accumulated in double
 * regardless of the storage form (the reference's `real` is double,
 * defs.h:77; sim_4.cu:257-300, state.hpp:1469-1492).  out[0..5] receives:
 *   [0] sum rho              (mass)
 *   [1] sum rho*|u|^2        (twice the kinetic energy)
 *   [2] count of sites with rho != rho (NaN check)
 *   [3] sum (rho-1)^2
 *   [4] sum |u|^2
 *   [5] number of sites summed
 * `out` is a HOST pointer; the call synchronises `stream`.                 */
int tnl_lbm_d3q27_reduce_macro(const tnl_lbm_config* cfg, const tnl_lbm_d3q27_data* SD,
                               double out[6], void* stream);

/* ---- FP64 reductions without allocation ------------------------------------
 * tnl_lbm_d3q27_reduce_macro with the same kernels and the same six results,
 * written to `out_device` (a DEVICE array of 6 doubles) asynchronously on
 * `stream`: no allocation, no copy to the host, no synchronisation.  No
 * counterpart in the reference: it is for callers that cannot allocate (code
 * under HIP graph capture, serving loops, harnesses that forbid hipMalloc).
 * Its partial sums live in a fixed device scratch of 1024 rows of 6 doubles
 * (48 KiB per GPU, held until the process ends); the reduction grid never
 * needs more, whatever the lattice size.  All calls on one GPU share that
 * scratch: calls that overlap on one GPU (two streams, two host threads)
 * race and must be ordered by the caller.                                   */
int tnl_lbm_d3q27_reduce_macro_noalloc(const tnl_lbm_config* cfg, const tnl_lbm_d3q27_data* SD, double* out_device,
                                       void* stream);

/* ---- FP64 velocity-gradient reductions -----------------------------------
 * The enstrophy diagnostic of sim_4.cu:282-300 (StateLocal::probe1): with the
 * velocity gradient G_ab = (v_a(r + e_b) - v_a(r - e_b)) / 2 of
 * computeGradVel (sim_4.cu:20-60), in lattice units, from the macro array:
 *   out[0] = sum rho |G|^2        (the reference divides by rho_0 2 V)
 *   out[1] = sum rho |curl v|^2   (computeVorticity, sim_4.cu:62-82)
 * Neighbours wrap periodically on an axis that is not distributed
 * (flags TNL_LBM_FLAG_DISTRIBUTED_*) and come from the overlap layer on one
 * that is.  `out` is a HOST pointer; the call synchronises `stream`.       */
int tnl_lbm_d3q27_reduce_gradients(const tnl_lbm_config* cfg, const tnl_lbm_d3q27_data* SD, uint32_t flags,
                                   double out[2], void* stream);

/* ---- halo exchange of an x-slab split ------------------------------------
 * One lattice split into x-slabs, each stored with a ghost plane per side
 * (ox >= 1) and stepped with TNL_LBM_FLAG_DISTRIBUTED_X.  `lower`'s +x face
 * meets `upper`'s -x face (lower == upper: a slab that is its own periodic
 * neighbour).  Every face moves the nine slots with c_x != 0 that cross it,
 * one contiguous (z, y) plane each:
 *   TNL_LBM_HALO_PULL  boundary planes -> the neighbour's ghost plane, for
 *                      the next step's pulls: before an A-B step (on df[0],
 *                      the array that step reads), or after an A-A even step
 *                      (then the slots hold opposite directions);
 *   TNL_LBM_HALO_PUSH  A-A only, after an odd step: the values that step
 *                      wrote into the ghost planes -> the neighbour's
 *                      boundary plane.
 * Replaces the per-slot DistributedNDArraySynchronizer passes of
 * lbm_block.hpp:434-470 (sync directions df_sync_directions, defs.h:352-385,
 * reversed on even A-A steps and shifted by one site on odd ones).  The
 * copies are asynchronous on `stream`; the two lattices may live on
 * different GPUs that have peer access.  Both must have the same Y, Z, oy,
 * oz and storage form.                                                      */
enum tnl_lbm_halo_phase { TNL_LBM_HALO_PULL = 0, TNL_LBM_HALO_PUSH = 1 };
int tnl_lbm_d3q27_halo_x(const tnl_lbm_config* cfg, const tnl_lbm_d3q27_data* lower,
                         const tnl_lbm_d3q27_data* upper, int phase, void* stream);

/* ---- adjoint of one step ------------------------------------------------
 * The transposed Jacobian of one tnl_lbm_d3q27_step with the default
 * cumulant and the A-B pattern, taken at the state SD->df[0] (the state
 * that step reads):
 *   lambda      <- (d df_out / d df[0])^T lambda_next
 *   grad_inflow += (d df_out / d inflow velocity)^T lambda_next
 * lambda_next / lambda: 27*XYZ elements of the storage type (F64 or F32),
 * same layout as the populations (F32: float fields, double arithmetic, so
 * the rounding of long adjoint runs stays that of the stored fields); lambda
 * is overwritten, must not alias
 * lambda_next or df[0].  grad_inflow: 3*Y*Z doubles (component-major, index
 * Y*z + y as the inflow profiles) accumulated, or NULL.  Geometry codes with
 * an adjoint: FLUID, PERIODIC, WALL, INFLOW, OUTFLOW_EQ, OUTFLOW_RIGHT,
 * NOTHING; other codes yield NaN at their sources, and so does an
 * OUTFLOW_RIGHT site off the faces of the box.  Single block (no overlaps).
 * The reference's adjoint machinery is for SRT
 * (d3q27/col_srt_adjoint.h:4-110, streaming_AB.h:246-281 streamingAdjoint,
 * bc.h:295-398, driven by sim_adjoint.cu:278-491); this entry point is the
 * same per-step operation for the default cumulant.                        */
int tnl_lbm_d3q27_adjoint_step(const tnl_lbm_config* cfg, const tnl_lbm_d3q27_data* SD,
                               const void* lambda_next, void* lambda, double* grad_inflow, void* stream);

/* ---- moment recording for the adjoint -------------------------------------
 * The Jacobian of the default cumulant step at a site depends on the pulled
 * populations only through their ten raw moments (m0, j, Pi).  A forward
 * step can record them, so that the adjoint reads 10 values per site instead
 * of pulling 27 populations, and a run need not be recomputed backwards.
 *
 * tnl_lbm_d3q27_step_record: the same A-B step as tnl_lbm_d3q27_step over
 * the whole local box (df[0] -> df[1], no macro output), with identical
 * results, that also writes into `tape` the moments m0, jx, jy, jz, Pxx,
 * Pyy, Pzz, Pxy, Pxz, Pyz (k = 0..9) of the populations pulled at each
 * FLUID, PERIODIC, OUTFLOW_RIGHT and OUTFLOW_EQ site s (other sites are left
 * untouched), at element ((s / 64) * 10 + k) * 64 + s % 64: blocks of 64
 * sites, ceil(XYZ / 64) * 640 elements in all.  tape_f64 = 1: double elements (F64 storage only; the adjoint
 * is then the same as from the state); 0: float elements.
 * tnl_lbm_d3q27_adjoint_step_recorded: tnl_lbm_d3q27_adjoint_step with the
 * Jacobian taken from a tape that tnl_lbm_d3q27_step_record wrote for that
 * step; SD->df is not read.  No counterpart in the reference: its adjoint
 * re-reads the stored primal macro fields (sim_adjoint.cu:189-241).        */
int tnl_lbm_d3q27_step_record(const tnl_lbm_config* cfg, const tnl_lbm_d3q27_data* SD, void* tape, int tape_f64,
                              void* stream);
/* ---- adjoint of the D3Q7 part of one coupled step ---------------------------
 * The passive scalar of tnl_lbm_nse_ade_step (A-B pattern) does not act on
 * the flow, and for a given velocity its update is linear in the D3Q7
 * populations, so its adjoint needs only the velocity the step used:
 *   lambda      <- (d g_out / d g_in)^T lambda_next
 *   grad_source += (d g_out / d S)^T lambda_next = sum_q w_q lambda_next_q at
 *                  FLUID, PERIODIC, SOLID and OUTFLOW_RIGHT sites
 * velocity: vx, vy, vz as 3 consecutive planes of XYZ elements of
 * cfg->macro_real, as the step wrote them in planes 1..3 of the NSE macro
 * array with TNL_LBM_FLAG_OUTPUT_MACRO.  lambda_next / lambda: 7*XYZ elements
 * of the compute type (double for F64 storage, float otherwise), layout of
 * the populations; lambda must not alias lambda_next.  grad_source: XYZ
 * doubles, accumulated, or NULL.  nse supplies sizes and the D3Q27 map (whose
 * PERIODIC code decides the pull's wrap, as in the coupled kernel); its
 * populations are not read.  D3Q7 codes with an adjoint: FLUID, PERIODIC,
 * SOLID, WALL, INFLOW, OUTFLOW_RIGHT (on the faces of the box), NOTHING;
 * others yield NaN.  Single block.  No counterpart in the reference (its
 * adjoint is for the D3Q27 SRT flow, sim_adjoint.cu); the forward step it
 * reverses is kernels.h:114-196 with d3q7/col_*.h and d3q7/bc.h.           */
int tnl_lbm_d3q7_adjoint_step(const tnl_lbm_config* cfg, const tnl_lbm_d3q27_data* nse, const tnl_lbm_d3q7_data* ade,
                              const void* velocity, const void* lambda_next, void* lambda, double* grad_source,
                              void* stream);
int tnl_lbm_d3q27_adjoint_step_recorded(const tnl_lbm_config* cfg, const tnl_lbm_d3q27_data* SD, const void* tape,
                                        int tape_f64, const void* lambda_next, void* lambda, double* grad_inflow,
                                        void* stream);

#ifdef __cplusplus
}
#endif
#endif /* TNL_LBM_AMD_H */
