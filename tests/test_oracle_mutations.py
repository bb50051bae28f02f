"""The oracle's pins catch plausible mistakes: each case plants one mistake
in a copy of oracle/lbm_oracle.cpp (a mirrored direction, a sign, a source
index, swapped weights, a wrong pull column, a wrong relaxation rate or
stabiliser),
builds it, and runs the pin that is meant to catch it against the mutant,
which must fail.  CPU only; nothing of the product is involved."""
import shutil
import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
SRC = ROOT / "oracle" / "lbm_oracle.cpp"

MUTANTS = {
    "sym_top_direction": ("case GEO_SYM_TOP: axis = 2; sgn = -1; break;",
                          "case GEO_SYM_TOP: axis = 2; sgn = 1; break;",
                          "test_oracle_bc.py", "symmetry"),
    "moment_inflow_rho": ("					KS.rho = (real) 1.0 / (1 - KS.vx)",
                          "					KS.rho = (real) 1.0 / (1 + KS.vx)",
                          "test_oracle_bc.py", "moment_inflow"),
    "bounce_back_inflow_source": ("f[ppm] = unstreamed(G, SD, mmp, x, y, z)", "f[ppm] = unstreamed(G, SD, mpm, x, y, z)",
                                  "test_oracle_bc.py", "bounce_back"),
    "interpolation_weights": (
        "f[i] = SpeedOfSound * unstreamed(G, SD, i, xm, yy, zz) + (1 - SpeedOfSound) * unstreamed(G, SD, i, x, yy, zz);",
        "f[i] = SpeedOfSound * unstreamed(G, SD, i, x, yy, zz) + (1 - SpeedOfSound) * unstreamed(G, SD, i, xm, yy, zz);",
        "test_oracle_bc.py", "interpolating"),
    "d3q7_symmetry": ("case A_SYM_TOP: KS.f[zzm] = KS.f[zzp];", "case A_SYM_TOP: KS.f[zzp] = KS.f[zzm];",
                      "test_oracle_bc.py", "d3q7"),
    "d3q7_outflow_pull": ("				xxp = xx = xm;", "				xxp = xm;", "test_oracle_bc.py", "outflow_right_pulls"),
    "kbc_stabiliser": ("const dreal gamma = no1 / beta - (no2 - no1 / beta) * num / den;",
                       "const dreal gamma = no1 / beta - (no2 - no1 / beta) * den / num;",
                       "test_oracle_alt.py", "stationary_point"),
    "kbc_bgk_limit": ("const dreal gamma = no1 / beta - (no2 - no1 / beta) * num / den;",
                      "const dreal gamma = no2 + 0 * num / den;", "test_oracle_alt.py", "stationary_point"),
    "cum2017_limiter": ("const dreal lambda3 = (dreal) (0.01), lambda4 = (dreal) (0.01), lambda5 = (dreal) (0.01);",
                        "const dreal lambda3 = (dreal) (0.01), lambda4 = (dreal) (0.01), lambda5 = (dreal) (0.01) / rho;",
                        "test_oracle_site.py", "limiter_form"),
    "relaxation_rate": ("const dreal omega1 = no1 / (no3 * KS.lbmViscosity + n1o2);",
                        "const dreal omega1 = no1 / (no3 * KS.lbmViscosity + no1);",
                        "test_oracle_lattice.py", "shear_wave"),
    # the 2017 rates and coefficients, and the default variant's product terms
    "cum2017_w3_on_w4_channel": ("const dreal omega120m102 = omega4 + (no1 - omega4) * std::fabs(C_120 - C_102)",
                                 "const dreal omega120m102 = omega3 + (no1 - omega3) * std::fabs(C_120 - C_102)",
                                 "test_oracle_cumulant_space.py", "closed_forms_per_channel"),
    "cum2017_w5_coefficient": ("omega1 * omega2 * (no18 - no13 * omega1)", "omega1 * omega2 * (no18 - (dreal) 12 * omega1)",
                               "test_oracle_cumulant_space.py", "closed_forms_per_channel"),
    "cum2017_A_coefficient": ("omega2 * omega2 * (omega1 * (no10 - no3 * omega1) - no4))",
                              "omega2 * omega2 * (omega1 * (no10 - no2 * omega1) - no4))",
                              "test_oracle_cumulant_space.py", "relaxation_rules"),
    "cum2017_B_coefficient": ("(no4 * omega1 * omega2 * (no9 * omega1 - no16)", "(no4 * omega1 * omega2 * (no8 * omega1 - no16)",
                              "test_oracle_cumulant_space.py", "relaxation_rules"),
    "cum2017_eq44_symmetrised": ("A * rho * (Dxu + Dyv - no2 * Dzw)", "A * rho * (Dyv + Dzw - no2 * Dxu)",
                                 "test_oracle_cumulant_space.py", "relaxation_rules"),
    "cum2017_eq83_index": ("Cs_122 + (ks_020 * ks_102 + ks_002 * ks_120 + no4 * ks_011 * ks_111",
                           "Cs_122 + (ks_020 * ks_102 + ks_002 * ks_120 + no4 * ks_101 * ks_111",
                           "test_oracle_cumulant_space.py", "relaxation_rules"),
    "site_force_plane": ("		fz += (real) F[2 * SD->XYZ + s];\n	}\n	else {",
                         "		fz += (real) F[SD->XYZ + s];\n	}\n	else {",
                         "test_oracle_forcing.py", "own_force"),
    "cum_default_cs110_squared": ("const dreal ks_220 = (ks_020 * ks_200 + no2 * ks_110 * ks_110) * rho_inv;",
                                  "const dreal ks_220 = (ks_020 * ks_200 + ks_110 * ks_110) * rho_inv;",
                                  "test_oracle_cumulant_space.py", "relaxation_rules"),
}


@pytest.fixture(scope="module")
def mutants(tmp_path_factory):
    """All mutants, compiled concurrently."""
    if not shutil.which("g++"):
        pytest.skip("g++ missing")
    d = tmp_path_factory.mktemp("mutants")
    src = SRC.read_text()
    procs = {}
    for name, (old, new, _, _) in MUTANTS.items():
        assert src.count(old) == 1, name
        cpp = d / f"{name}.cpp"
        cpp.write_text(src.replace(old, new))
        cmd = ["g++", "-O0", "-std=c++17", "-fPIC", "-fopenmp", "-ffp-contract=off", "-iquote", str(ROOT / "oracle"),
               "-shared", "-o", str(d / f"{name}.so"), str(cpp)]
        procs[name] = subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True)
    for name, p in procs.items():
        _, err = p.communicate()
        assert p.returncode == 0, err[-2000:]
    return d


@pytest.mark.parametrize("name", list(MUTANTS))
def test_pin_catches_planted_mistake(mutants, tmp_path, name):
    _, _, test_file, selector = MUTANTS[name]
    so = mutants / f"{name}.so"
    code = (f"import sys; sys.path[:0] = [{str(ROOT / 'tnl-lbm_amd')!r}, {str(ROOT / 'oracle')!r}, "
            f"{str(ROOT / 'tests')!r}]\n"
            "import oracle; from pathlib import Path\n"
            f"oracle.LIB_PATH = Path({str(so)!r})\n"
            "import pytest\n"
            f"sys.exit(pytest.main(['-q', '-x', '-k', {selector!r}, {str(ROOT / 'tests' / test_file)!r}, "
            "'-p', 'no:cacheprovider', '-m', 'not gpu']))\n")
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, cwd=tmp_path)
    assert r.returncode == 1 and "failed" in r.stdout, (name, r.returncode, r.stdout[-800:])
