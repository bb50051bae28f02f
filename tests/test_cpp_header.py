"""include/tnl_lbm_amd.hpp maps the reference's plugin types onto the C ABI.
Compiled here with g++ against stand-ins that carry the reference's own
template parameter lists (lbm3d/d3q27/col_*.h, lbm3d/d3q7/col_*.h, declared
after the header, with their default arguments), the LBM_Data / NSE_Data /
ADE_Data field names and types (lbm_data.h:8-261: dreal* dfs[DFMAX] with
DFMAX from defs.h:45-66, dreal-typed inflow profiles) for TraitsDP and
TraitsSP: the configuration and data structs it builds are checked for each
compile-time switch the reference uses (USE_GEIER_CUM_2017,
USE_GEIER_CUM_ANTIALIAS, AA_PATTERN).  Under AA_PATTERN dfs has one entry
and df[1] must come out NULL: a read past dfs[0] would pick up the next
member (dmacro) there.  An operator without a counterpart must fail to
compile."""
import shutil
import subprocess
from pathlib import Path

import pytest

from pytnl_lbm import _abi

ROOT = Path(__file__).resolve().parents[1]
LIBDIR = ROOT / "tnl-lbm_amd" / "pytnl_lbm"

SRC = r"""
#include <tnl_lbm_amd.hpp>
#include <cstdio>

// DFMAX as the reference's defs.h:45-66 declares it
#ifdef AA_PATTERN
enum : unsigned char { df_cur, DFMAX };
#else
enum : unsigned char { df_cur, df_out, DFMAX };
#endif

template <typename TRAITS> struct D3Q27_EQ {};
template <typename TRAITS> struct D3Q27_EQ_ENTROPIC {};
template <typename TRAITS> struct D3Q7_EQ {};
template <typename TRAITS, typename LBM_EQ = D3Q27_EQ<TRAITS>> struct D3Q27_CUM {};
template <typename TRAITS, typename LBM_EQ = D3Q27_EQ<TRAITS>> struct D3Q27_MRT {};
template <typename TRAITS, typename LBM_EQ = D3Q27_EQ_ENTROPIC<TRAITS>> struct D3Q27_KBC_N4 {};
template <typename TRAITS, typename LBM_EQ = D3Q27_EQ<TRAITS>> struct D3Q27_KBC_C1 {};
template <typename TRAITS, typename LBM_EQ = D3Q27_EQ<TRAITS>> struct D3Q27_SRT {};
template <typename TRAITS, typename LBM_EQ = D3Q7_EQ<TRAITS>> struct D3Q7_SRT {};
template <typename TRAITS, typename LBM_EQ = D3Q7_EQ<TRAITS>> struct D3Q7_MRT {};
template <typename TRAITS, typename LBM_EQ = D3Q7_EQ<TRAITS>> struct D3Q7_CLBM {};

template <typename D> struct Traits { using dreal = D; };
using TraitsDP = Traits<double>;
using TraitsSP = Traits<float>;
template <typename T, typename C> struct NSE { using TRAITS = T; using COLL = C; };

struct Indexer { template <int d> long getOverlap() const { return d + 1; } };
struct Idx3 { long a, b, c; long x() const { return a; } long y() const { return b; } long z() const { return c; } };
struct Dist { bool a, b, c; bool x() const { return a; } bool y() const { return b; } bool z() const { return c; } };
static double f0[10], f1[10], m[10];
static float s0[10], s1[10], sm[10], sd[10];
static short mp[10], amp[10];
static double px[4], py[4], pz[4];
static float qx[4], qy[4], qz[4];
// LBM_Data<TRAITS> (lbm_data.h:8-60): dreal* dfs[DFMAX]
template <typename R>
struct Base {
    bool even_iter = false; Indexer indexer; long XYZ = 777; R lbmViscosity = (R) 0.0125;
    R* dfs[DFMAX]; R* dmacro; short* dmap = mp;
    long X() { return 5; } long Y() { return 6; } long Z() { return 7; }
};
struct ConstInflow : Base<double> {
    double fx = 1, fy = 2, fz = 3; double inflow_vx = 0.1, inflow_vy = 0.2, inflow_vz = 0.3;
    ConstInflow() { dfs[0] = f0; if (DFMAX > 1) dfs[DFMAX - 1] = f1; dmacro = m; }
};
struct Profile : Base<double> {
    double fx = 0, fy = 0, fz = 0; double* inflow_vx = px; double* inflow_vy = py; double* inflow_vz = pz;
    Profile() { dfs[0] = f0; if (DFMAX > 1) dfs[DFMAX - 1] = f1; dmacro = m; }
};
// NSE_Data_InflowProfile<TraitsSP>: float profiles (lbm_data.h:100-104)
struct ProfileSP : Base<float> {
    float fx = 0, fy = 0, fz = 0; float* inflow_vx = qx; float* inflow_vy = qy; float* inflow_vz = nullptr;
    ProfileSP() { dfs[0] = s0; if (DFMAX > 1) dfs[DFMAX - 1] = s1; dmacro = sm; }
};
// ADE_Data_ConstInflow<TraitsSP> (lbm_data.h:212-261)
struct AdeSP : Base<float> {
    float* diffusion_coefficient_ptr = sd; bool* phi_transfer_direction_ptr = nullptr; float phiTransferCoefficient = 0;
    float inflow_phi = 0.75f;
    AdeSP() { dfs[0] = s1; if (DFMAX > 1) dfs[DFMAX - 1] = s0; dmacro = sm + 1; dmap = amp; lbmViscosity = 0.0125f; }
};
// ADE_Data<TraitsSP> without an inflow value
struct AdePlainSP : Base<float> {
    float* diffusion_coefficient_ptr = nullptr;
    AdePlainSP() { dfs[0] = s1; if (DFMAX > 1) dfs[DFMAX - 1] = s0; dmacro = sm; }
};

// a MACRO with the IBM force quantities (d3q27/macro.h:15-25)
struct MacroIBM { enum : unsigned char { N, e_rho, e_vx, e_vy, e_vz, e_fx, e_fy, e_fz }; };
static double big[8 * 777];
struct WithForce : Base<double> { WithForce() { dfs[0] = f0; dmacro = big; } };

template <typename N, typename DATA>
int unused_launch(DATA& d) { return tnl_lbm_amd::step<N>(d, Idx3{0, 0, 0}, Idx3{5, 6, 7}, Dist{false, false, false}, true); }
template <typename N, typename A, typename D1, typename D2>
int unused_coupled(D1& n, D2& a) { return tnl_lbm_amd::step_coupled<N, A>(n, a, Idx3{0, 0, 0}, Idx3{5, 6, 7}, Dist{false, false, false}, true); }
template <typename N, typename DATA>
int unused_lists(DATA& d, tnl_lbm_amd::row_lists& L)
{
    if (int rc = tnl_lbm_amd::classify_rows(d, L))
        return rc;
    return tnl_lbm_amd::step<N>(d, Idx3{0, 0, 0}, Idx3{5, 6, 7}, Dist{false, false, false}, true, nullptr, -1, &L);
}
template <typename N, typename A, typename D1, typename D2>
int unused_coupled_lists(D1& n, D2& a, tnl_lbm_amd::row_lists& L)
{
    if (int rc = tnl_lbm_amd::classify_rows_coupled<A>(n, a, L))
        return rc;
    return tnl_lbm_amd::step_coupled<N, A>(n, a, Idx3{0, 0, 0}, Idx3{5, 6, 7}, Dist{false, false, false}, true,
                                           nullptr, -1, 0.0, 1.0, &L);
}

int main()
{
    using namespace tnl_lbm_amd;
    auto c1 = config_for<NSE<TraitsDP, D3Q27_CUM<TraitsDP>>>();
    auto c2 = config_for<NSE<TraitsSP, D3Q27_MRT<TraitsSP>>>();
    auto c3 = config_for<NSE<TraitsDP, D3Q27_KBC_N4<TraitsDP>>>();
    auto c4 = config_for<NSE<TraitsSP, D3Q27_KBC_C1<TraitsSP>>>(TNL_LBM_STORE_I16_DEV);
    std::printf("cfg %d %d %d %d | %d %d %d | %d | %d %d | %u %u\n", c1.collision, c1.storage, c1.pattern, c1.macro_real,
                c2.collision, c2.storage, c2.macro_real, c3.collision, c4.collision, c4.storage, c1.struct_size, c1.version);
    ConstInflow a; Profile b; ProfileSP c;
    auto s = data_for(a);
    auto t = data_for(b);
    auto u = data_for(c);
    std::printf("data %ld %ld %ld %ld %ld %ld %ld %d %g %g %g %g %d %d %d %g %g %g %d\n", (long) s.X, (long) s.Y, (long) s.Z,
                (long) s.ox, (long) s.oy, (long) s.oz, (long) s.XYZ, s.even_iter, s.lbm_viscosity, s.fx, s.fy, s.fz,
                s.df[0] == f0, s.df[1] == (DFMAX > 1 ? (void*) f1 : nullptr), s.map == mp, s.inflow_vx, s.inflow_vy,
                s.inflow_vz, s.inflow_profile[0] == nullptr);
    std::printf("profile %d %d %d %g %d\n", t.inflow_profile[0] == px, t.inflow_profile[1] == py, t.inflow_profile[2] == pz,
                t.inflow_vx, t.inflow_profile_real);
    std::printf("profile_sp %d %d %d %d %d %d\n", u.inflow_profile[0] == qx, u.inflow_profile[1] == qy,
                u.inflow_profile[2] == nullptr, u.inflow_profile_real, u.df[0] == s0,
                u.df[1] == (DFMAX > 1 ? (void*) s1 : nullptr));
    AdeSP e; AdePlainSP e2;
    auto ad = ade_data_for<D3Q7_CLBM<TraitsSP>>(e, 0.25, 1000.0);
    auto ad2 = ade_data_for<D3Q7_MRT<TraitsSP>>(e2);
    std::printf("ade %ld %ld %ld %ld %ld %ld %ld %d %d %g %d %d %d %d %d %d %g %g %g | %d %g %d\n", (long) ad.X, (long) ad.Y,
                (long) ad.Z, (long) ad.ox, (long) ad.oy, (long) ad.oz, (long) ad.XYZ, ad.even_iter, ad.collision,
                ad.lbm_diffusivity, ad.df[0] == s1, ad.df[1] == (DFMAX > 1 ? (void*) s0 : nullptr), ad.macro == sm + 1,
                ad.map == amp, ad.diffusion_coefficient == sd, ad.source == nullptr, ad.inflow_phi, ad.phi_ref,
                ad.store_scale, ad2.collision, ad2.inflow_phi, ad2.diffusion_coefficient == nullptr);
    std::printf("flags %u %u\n", flags_for(Dist{true, false, true}, true), flags_for(Dist{false, true, false}, false));
    WithForce wf;
    std::printf("force %d\n", macro_force_field<MacroIBM>(wf) == (const void*) (big + 5 * 777));
    if (false) {
        unused_launch<NSE<TraitsDP, D3Q27_CUM<TraitsDP>>>(a);
        unused_coupled<NSE<TraitsSP, D3Q27_CUM<TraitsSP>>, NSE<TraitsSP, D3Q7_SRT<TraitsSP>>>(c, e);
        row_lists L;
        unused_lists<NSE<TraitsDP, D3Q27_CUM<TraitsDP>>>(a, L);
        unused_coupled_lists<NSE<TraitsSP, D3Q27_CUM<TraitsSP>>, NSE<TraitsSP, D3Q7_SRT<TraitsSP>>>(c, e, L);
    }
    return 0;
}
"""


def build(tmp_path, defines=(), extra=""):
    if not shutil.which("g++"):
        pytest.skip("g++ missing")
    src = tmp_path / "t.cpp"
    src.write_text(SRC + extra)
    exe = tmp_path / "t"
    cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-Werror=array-bounds", *[f"-D{d}" for d in defines], f"-I{ROOT / 'include'}", str(src), "-o", str(exe),
           f"-L{LIBDIR}", "-ltnl_lbm_amd", f"-Wl,-rpath,{LIBDIR}", "-Wl,--allow-shlib-undefined"]
    return subprocess.run(cmd, capture_output=True, text=True), exe


@pytest.mark.parametrize("defines,cum,pattern", [((), 0, 1), (("USE_GEIER_CUM_2017",), 1, 1),
                                                 (("USE_GEIER_CUM_2017", "USE_GEIER_CUM_ANTIALIAS"), 2, 1),
                                                 (("AA_PATTERN",), 0, 0)])
def test_header_maps_reference_types(tmp_path, defines, cum, pattern):
    if not (LIBDIR / "libtnl_lbm_amd.so").exists():
        pytest.skip("library not built")
    r, exe = build(tmp_path, defines)
    assert r.returncode == 0, r.stderr
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split("\n")
    cfg = out[0].split()
    # CUM variant, F64, pattern, F64 macro | MRT_LES, F32, F32 macro | KBC_N4 | KBC_C1, I16_DEV
    assert cfg[1:5] == [str(cum), "0", str(pattern), "0"]
    assert cfg[6:9] == ["3", "1", "1"]
    assert cfg[10] == "7" and cfg[12:14] == ["8", "5"]
    assert cfg[16] == str(_abi.ABI_VERSION)  # ABI version
    aa = "AA_PATTERN" in defines
    data = out[1].split()[1:]
    assert data[:7] == ["5", "6", "7", "1", "2", "3", "777"]
    assert data[7:12] == ["0", "0.0125", "1", "2", "3"]
    assert data[12:15] == ["1", "1", "1"]  # df[1] is dfs[1] under A-B, NULL under A-A (DFMAX = 1)
    assert data[15:19] == ["0.1", "0.2", "0.3", "1"]
    assert out[2].split()[1:] == ["1", "1", "1", "0", "0"]  # double profiles: inflow_profile_real = F64
    assert out[3].split()[1:] == ["1", "1", "1", "1", "1", "1"]  # float profiles: F32
    ade = out[4].split()[1:]
    assert ade[:8] == ["5", "6", "7", "1", "2", "3", "777", "0"]
    assert ade[8:10] == ["2", "0.0125"]  # D3Q7_CLBM, ADE_Data::lbmViscosity
    assert ade[10:16] == ["1", "1", "1", "1", "1", "1"]
    assert ade[16:19] == ["0.75", "0.25", "1000"]
    assert ade[20:23] == ["1", "1", "1"]  # D3Q7_MRT; no inflow_phi member -> 1; no per-site D
    assert out[5].split()[1:] == [str(1 | 2 | 8), str(4)]
    assert out[6].split()[1:] == ["1"]  # the force planes e_fx..e_fz of the macro array
    assert aa == (cfg[3] == "0")


def test_operator_without_counterpart_does_not_compile(tmp_path):
    r, _ = build(tmp_path, extra="\nauto bad = tnl_lbm_amd::config_for<NSE<TraitsDP, D3Q27_SRT<TraitsDP>>>();\n")
    assert r.returncode != 0
    assert "no counterpart" in r.stderr
