/* C++ face of libtnl_lbm_amd for code built on the reference's plugin API.
 *
 * The reference selects its lattice update at compile time through
 * LBM_CONFIG<TRAITS, KernelStruct, DATA, COLL, EQ, STREAMING, BC, MACRO>
 * (lbm3d/defs.h:288-310) and launches cudaLBMKernel<NSE> per block
 * (lbm3d/state.hpp:1302-1309).  This header maps those same plugin types
 * onto the run-time tnl_lbm_config of the C ABI, so that a reference-style
 * State::SimUpdate keeps its types and replaces only the launch:
 *
 *     #include <tnl_lbm_amd.hpp>
 *     ...
 *     tnl_lbm_amd::step<NSE>(block.data, offset, end, block.is_distributed(), compute_macro, stream);
 *
 * and the coupled update of State_NSE_ADE (state_NSE_ADE.h:201-210) the same
 * way:
 *
 *     tnl_lbm_amd::step_coupled<NSE, ADE>(block_nse.data, block_ade.data, offset, end,
 *                                         block_nse.is_distributed(), compute_macro, stream);
 *
 * Mapped types (forward-declared below with the reference's own template
 * parameter lists, lbm3d/d3q27/col_*.h; declared without default arguments so
 * that the reference's headers may come before or after this one):
 *   D3Q27_CUM<T, EQ>                 -> TNL_LBM_COLL_CUM, or CUM_2017 /
 *                                       CUM_2017_ANTIALIAS under the
 *                                       reference's USE_GEIER_CUM_2017 /
 *                                       USE_GEIER_CUM_ANTIALIAS (defs.h:314-315)
 *   D3Q27_MRT<T, EQ>                 -> TNL_LBM_COLL_MRT_LES
 *   D3Q27_KBC_N1..N4, D3Q27_KBC_C1..C4 -> TNL_LBM_COLL_KBC_*
 *   D3Q27_STREAMING                  -> A-A under AA_PATTERN, else A-B
 *                                       (core.h:20-25)
 *   TRAITS::dreal double / float     -> F64 / F32 storage, macro and inflow
 *                                       profile arrays
 *   D3Q7_SRT / D3Q7_MRT / D3Q7_CLBM  -> TNL_LBM_ADE_COLL_SRT / MRT / CLBM
 *                                       (d3q7/col_srt.h, col_mrt.h, col_clbm.h)
 * LBM_Data::dfs has DFMAX entries (defs.h:45-66): one under AA_PATTERN, and
 * dfs[1] (df_out) is read only where the array has a second entry.
 * The deviation storage forms (F32_DEV ... I16_DEV) have no reference type:
 * pass the storage explicitly.  An unsupported collision fails to compile.
 */
#pragma once

#include <cstdint>
#include <type_traits>

#include "tnl_lbm_amd.h"

template <typename TRAITS, typename LBM_EQ>
struct D3Q27_CUM;
template <typename TRAITS, typename LBM_EQ>
struct D3Q27_MRT;
template <typename TRAITS, typename LBM_EQ>
struct D3Q27_KBC_N1;
template <typename TRAITS, typename LBM_EQ>
struct D3Q27_KBC_N2;
template <typename TRAITS, typename LBM_EQ>
struct D3Q27_KBC_N3;
template <typename TRAITS, typename LBM_EQ>
struct D3Q27_KBC_N4;
template <typename TRAITS, typename LBM_EQ>
struct D3Q27_KBC_C1;
template <typename TRAITS, typename LBM_EQ>
struct D3Q27_KBC_C2;
template <typename TRAITS, typename LBM_EQ>
struct D3Q27_KBC_C3;
template <typename TRAITS, typename LBM_EQ>
struct D3Q27_KBC_C4;
template <typename TRAITS, typename LBM_EQ>
struct D3Q7_SRT;
template <typename TRAITS, typename LBM_EQ>
struct D3Q7_MRT;
template <typename TRAITS, typename LBM_EQ>
struct D3Q7_CLBM;

namespace tnl_lbm_amd {

template <typename COLL>
struct collision_id
{
	static_assert(sizeof(COLL) == 0, "this collision operator has no counterpart in libtnl_lbm_amd");
};

template <typename T, typename E>
struct collision_id<D3Q27_CUM<T, E>>
{
#if defined(USE_GEIER_CUM_2017) && defined(USE_GEIER_CUM_ANTIALIAS)
	static constexpr int value = TNL_LBM_COLL_CUM_2017_ANTIALIAS;
#elif defined(USE_GEIER_CUM_2017)
	static constexpr int value = TNL_LBM_COLL_CUM_2017;
#elif defined(USE_GEIER_CUM_ANTIALIAS)
	static_assert(sizeof(T) == 0, "the antialias correction is built together with the 2017 limiter only");
#else
	static constexpr int value = TNL_LBM_COLL_CUM;
#endif
};
// clang-format off
template <typename T, typename E> struct collision_id<D3Q27_MRT<T, E>>    { static constexpr int value = TNL_LBM_COLL_MRT_LES; };
template <typename T, typename E> struct collision_id<D3Q27_KBC_N1<T, E>> { static constexpr int value = TNL_LBM_COLL_KBC_N1; };
template <typename T, typename E> struct collision_id<D3Q27_KBC_N2<T, E>> { static constexpr int value = TNL_LBM_COLL_KBC_N2; };
template <typename T, typename E> struct collision_id<D3Q27_KBC_N3<T, E>> { static constexpr int value = TNL_LBM_COLL_KBC_N3; };
template <typename T, typename E> struct collision_id<D3Q27_KBC_N4<T, E>> { static constexpr int value = TNL_LBM_COLL_KBC_N4; };
template <typename T, typename E> struct collision_id<D3Q27_KBC_C1<T, E>> { static constexpr int value = TNL_LBM_COLL_KBC_C1; };
template <typename T, typename E> struct collision_id<D3Q27_KBC_C2<T, E>> { static constexpr int value = TNL_LBM_COLL_KBC_C2; };
template <typename T, typename E> struct collision_id<D3Q27_KBC_C3<T, E>> { static constexpr int value = TNL_LBM_COLL_KBC_C3; };
template <typename T, typename E> struct collision_id<D3Q27_KBC_C4<T, E>> { static constexpr int value = TNL_LBM_COLL_KBC_C4; };
// clang-format on

// the D3Q7 operator of an ADE configuration (ADE::COLL)
template <typename COLL>
struct ade_collision_id
{
	static_assert(sizeof(COLL) == 0, "this D3Q7 collision operator has no counterpart in libtnl_lbm_amd");
};
// clang-format off
template <typename T, typename E> struct ade_collision_id<D3Q7_SRT<T, E>>  { static constexpr int value = TNL_LBM_ADE_COLL_SRT; };
template <typename T, typename E> struct ade_collision_id<D3Q7_MRT<T, E>>  { static constexpr int value = TNL_LBM_ADE_COLL_MRT; };
template <typename T, typename E> struct ade_collision_id<D3Q7_CLBM<T, E>> { static constexpr int value = TNL_LBM_ADE_COLL_CLBM; };
// clang-format on

// the streaming plugin the reference compiled in (core.h:20-25)
#ifdef AA_PATTERN
constexpr int pattern_id = TNL_LBM_PATTERN_AA;
#else
constexpr int pattern_id = TNL_LBM_PATTERN_AB;
#endif

template <typename DREAL>
constexpr int storage_of()
{
	static_assert(std::is_same<DREAL, double>::value || std::is_same<DREAL, float>::value, "dreal must be double or float");
	return std::is_same<DREAL, double>::value ? TNL_LBM_STORE_F64 : TNL_LBM_STORE_F32;
}

// tnl_lbm_config of an LBM_CONFIG bundle (NSE::COLL, NSE::TRAITS::dreal); storage < 0: the dreal form
template <typename NSE>
inline tnl_lbm_config config_for(int storage = -1)
{
	using dreal = typename NSE::TRAITS::dreal;
	tnl_lbm_config c{};
	c.struct_size = sizeof c;
	c.version = TNL_LBM_ABI_VERSION;
	c.collision = collision_id<typename NSE::COLL>::value;
	c.storage = storage < 0 ? storage_of<dreal>() : storage;
	c.pattern = pattern_id;
	c.macro_real = std::is_same<dreal, double>::value ? TNL_LBM_REAL_F64 : TNL_LBM_REAL_F32;
	return c;
}

namespace detail {
template <typename D, typename = void>
struct has_inflow : std::false_type
{};
template <typename D>
struct has_inflow<D, std::void_t<decltype(std::declval<D&>().inflow_vx)>> : std::true_type
{};
template <typename D, typename = void>
struct has_force : std::false_type
{};
template <typename D>
struct has_force<D, std::void_t<decltype(std::declval<D&>().fx)>> : std::true_type
{};
template <typename D, typename = void>
struct has_inflow_phi : std::false_type
{};
template <typename D>
struct has_inflow_phi<D, std::void_t<decltype(std::declval<D&>().inflow_phi)>> : std::true_type
{};

// dfs[0] and, where LBM_Data::dfs has a second entry (DFMAX > 1, defs.h:45-66), dfs[1]
template <typename DATA>
inline void copy_dfs(DATA& D, void* df[2])
{
	using arr = std::remove_reference_t<decltype(D.dfs)>;
	df[0] = D.dfs[0];
	if constexpr (std::extent<arr>::value > 1)
		df[1] = D.dfs[1];
	else
		df[1] = nullptr;
}
}  // namespace detail

// tnl_lbm_d3q27_data of a block's NSE::DATA (LBM_Data, NSE_Data,
// NSE_Data_ConstInflow or NSE_Data_InflowProfile, lbm_data.h:8-133)
template <typename DATA>
inline tnl_lbm_d3q27_data data_for(DATA& D)
{
	tnl_lbm_d3q27_data sd{};
	sd.struct_size = sizeof sd;
	sd.version = TNL_LBM_ABI_VERSION;
	sd.X = D.X();
	sd.Y = D.Y();
	sd.Z = D.Z();
	sd.ox = D.indexer.template getOverlap<0>();  // as kernels.h:39-41 reads them
	sd.oy = D.indexer.template getOverlap<1>();
	sd.oz = D.indexer.template getOverlap<2>();
	sd.XYZ = D.XYZ;
	sd.even_iter = D.even_iter ? 1 : 0;
	sd.lbm_viscosity = D.lbmViscosity;
	if constexpr (detail::has_force<DATA>::value) {
		sd.fx = D.fx;
		sd.fy = D.fy;
		sd.fz = D.fz;
	}
	detail::copy_dfs(D, sd.df);
	sd.macro = D.dmacro;
	sd.map = D.dmap;
	if constexpr (detail::has_inflow<DATA>::value) {
		using V = std::remove_reference_t<decltype(D.inflow_vx)>;
		if constexpr (std::is_pointer<V>::value) {
			// NSE_Data_InflowProfile: per-(y, z) arrays of dreal indexed Y*z + y
			using E = std::remove_cv_t<std::remove_pointer_t<V>>;
			static_assert(std::is_same<E, double>::value || std::is_same<E, float>::value,
						  "inflow profiles must be double or float arrays");
			sd.inflow_profile_real = std::is_same<E, float>::value ? TNL_LBM_REAL_F32 : TNL_LBM_REAL_F64;
			sd.inflow_profile[0] = D.inflow_vx;
			sd.inflow_profile[1] = D.inflow_vy;
			sd.inflow_profile[2] = D.inflow_vz;
		}
		else {
			sd.inflow_vx = D.inflow_vx;
			sd.inflow_vy = D.inflow_vy;
			sd.inflow_vz = D.inflow_vz;
		}
	}
	return sd;
}

// tnl_lbm_d3q7_data of a block's ADE::DATA (ADE_Data or ADE_Data_ConstInflow,
// lbm_data.h:212-261) for the D3Q7 operator ADE_COLL.  The reference has no
// source term (a TODO at lbm_data.h:229): `source` stays NULL.  phi_ref and
// store_scale matter for the deviation storage forms only.
template <typename ADE_COLL, typename DATA>
inline tnl_lbm_d3q7_data ade_data_for(DATA& D, double phi_ref = 0.0, double store_scale = 1.0)
{
	tnl_lbm_d3q7_data ad{};
	ad.struct_size = sizeof ad;
	ad.version = TNL_LBM_ABI_VERSION;
	ad.X = D.X();
	ad.Y = D.Y();
	ad.Z = D.Z();
	ad.ox = D.indexer.template getOverlap<0>();
	ad.oy = D.indexer.template getOverlap<1>();
	ad.oz = D.indexer.template getOverlap<2>();
	ad.XYZ = D.XYZ;
	ad.even_iter = D.even_iter ? 1 : 0;
	ad.collision = ade_collision_id<ADE_COLL>::value;
	ad.lbm_diffusivity = D.lbmViscosity;  // ADE_Data::diffusionCoefficient (lbm_data.h:233-242)
	detail::copy_dfs(D, ad.df);
	ad.macro = D.dmacro;
	ad.map = D.dmap;
	ad.diffusion_coefficient = D.diffusion_coefficient_ptr;
	ad.source = nullptr;
	if constexpr (detail::has_inflow_phi<DATA>::value)
		ad.inflow_phi = D.inflow_phi;
	else
		ad.inflow_phi = 1.0;  // ADE_Data_ConstInflow's default (lbm_data.h:252)
	ad.phi_ref = phi_ref;
	ad.store_scale = store_scale;
	return ad;
}

// The per-site force of a MACRO with force quantities e_fx, e_fy, e_fz (the
// IBM forcing of D3Q27_MACRO_Base, d3q27/macro.h:15-25, and the
// non-Newtonian MACRO, nonNewtonian.h:575-625): those macro planes are the
// three planes tnl_lbm_d3q27_data::force_field expects, so set
//     sd.force_field = tnl_lbm_amd::macro_force_field<NSE::MACRO>(block.data);
template <typename MACRO, typename DATA>
inline const void* macro_force_field(DATA& D)
{
	static_assert(MACRO::e_fy == MACRO::e_fx + 1 && MACRO::e_fz == MACRO::e_fx + 2,
				  "the force quantities must be consecutive macro planes");
	return D.dmacro + (long) MACRO::e_fx * D.XYZ;
}

// flags of the reference's call: distributed (x, y, z) and output_macro
template <typename DIST>
inline uint32_t flags_for(const DIST& distributed, bool output_macro)
{
	uint32_t f = output_macro ? TNL_LBM_FLAG_OUTPUT_MACRO : 0u;
	if (distributed.x())
		f |= TNL_LBM_FLAG_DISTRIBUTED_X;
	if (distributed.y())
		f |= TNL_LBM_FLAG_DISTRIBUTED_Y;
	if (distributed.z())
		f |= TNL_LBM_FLAG_DISTRIBUTED_Z;
	return f;
}

// Row lists (no counterpart in the reference): a caller-owned device array
// of local X*Z entries (a TNL array, say) and the three counts that
// tnl_lbm_classify_rows writes.  Classify once after the map is set up and
// again after every change of it, then pass the lists to step /
// step_coupled: whole-box steps of the deviation forms, and every whole-box
// coupled step, then run the rows of FLUID/PERIODIC (and WALL) sites
// through kernels without the other boundary codes.
struct row_lists
{
	uint32_t* rows = nullptr;
	int64_t count[3] = {0, 0, 0};
};

template <typename DATA>
inline int classify_rows(DATA& SD, row_lists& L, void* stream = nullptr)
{
	const tnl_lbm_d3q27_data sd = data_for(SD);
	return tnl_lbm_classify_rows(&sd, nullptr, L.rows, L.count, stream);
}

// the coupled step's lists, from both maps
template <typename ADE, typename NDATA, typename ADATA>
inline int classify_rows_coupled(NDATA& nse, ADATA& ade, row_lists& L, void* stream = nullptr)
{
	const tnl_lbm_d3q27_data sd = data_for(nse);
	const tnl_lbm_d3q7_data ad = ade_data_for<typename ADE::COLL>(ade);
	return tnl_lbm_classify_rows(&sd, &ad, L.rows, L.count, stream);
}

// One launch of cudaLBMKernel<NSE>(SD, offset, end, distributed, output_macro)
// (kernels.h:59-112): returns the library's status (0 = ok, negative =
// error; tnl_lbm_last_error() says which).  IDX3D: anything with x(), y(), z().
// lists: row lists of SD's map (classify_rows), or null.
template <typename NSE, typename DATA, typename IDX3D, typename DIST>
inline int step(DATA& SD, const IDX3D& offset, const IDX3D& end, const DIST& distributed, bool output_macro,
				void* stream = nullptr, int storage = -1, const row_lists* lists = nullptr)
{
	const tnl_lbm_config cfg = config_for<NSE>(storage);
	tnl_lbm_d3q27_data sd = data_for(SD);
	if (lists) {
		sd.rows = lists->rows;
		for (int c = 0; c < 3; c++)
			sd.row_count[c] = lists->count[c];
	}
	const int64_t b[3] = {(int64_t) offset.x(), (int64_t) offset.y(), (int64_t) offset.z()};
	const int64_t e[3] = {(int64_t) end.x(), (int64_t) end.y(), (int64_t) end.z()};
	return tnl_lbm_d3q27_step(&cfg, &sd, b, e, flags_for(distributed, output_macro), stream);
}

// One launch of cudaLBMKernel<NSE, ADE>(nse, ade, offset, end, distributed,
// output_macro) (kernels.h:114-196, launched at state_NSE_ADE.h:201-210):
// the D3Q27 update and the D3Q7 update at the same site.  Both configurations
// must share the dreal type; the storage form is the NSE one.
// lists: row lists of both maps (classify_rows_coupled), or null.
template <typename NSE, typename ADE, typename NDATA, typename ADATA, typename IDX3D, typename DIST>
inline int step_coupled(NDATA& nse, ADATA& ade, const IDX3D& offset, const IDX3D& end, const DIST& distributed,
						bool output_macro, void* stream = nullptr, int storage = -1, double phi_ref = 0.0,
						double store_scale = 1.0, const row_lists* lists = nullptr)
{
	static_assert(std::is_same<typename NSE::TRAITS::dreal, typename ADE::TRAITS::dreal>::value,
				  "the coupled kernel takes NSE and ADE arrays of one element type");
	const tnl_lbm_config cfg = config_for<NSE>(storage);
	const tnl_lbm_d3q27_data sd = data_for(nse);
	tnl_lbm_d3q7_data ad = ade_data_for<typename ADE::COLL>(ade, phi_ref, store_scale);
	if (lists) {
		ad.rows = lists->rows;
		for (int c = 0; c < 3; c++)
			ad.row_count[c] = lists->count[c];
	}
	const int64_t b[3] = {(int64_t) offset.x(), (int64_t) offset.y(), (int64_t) offset.z()};
	const int64_t e[3] = {(int64_t) end.x(), (int64_t) end.y(), (int64_t) end.z()};
	return tnl_lbm_nse_ade_step(&cfg, &sd, &ad, b, e, flags_for(distributed, output_macro), stream);
}

}  // namespace tnl_lbm_amd
