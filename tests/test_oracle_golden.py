"""The oracle reproduces the committed whole-step fixtures of tests/golden/
(written by tests/golden/make_fixtures.py) exactly: a regression pin of the
restatement for every operator, pattern, storage form and boundary set they
cover."""
from pathlib import Path

import numpy as np
import pytest

import oracle as O
from pytnl_lbm import _abi

GOLD = Path(__file__).resolve().parent / "golden"
FILES = sorted(GOLD.glob("*.npz"))


def run_oracle(d):
    X, Y, Z = d["f0"].shape[1:]
    geo = d["geo"] if "geo" in d else None
    st = int(d["storage"])
    h = O.HostLattice(X, Y, Z, collision=int(d["collision"]), storage=st, pattern=int(d["pattern"]),
                      nu=float(d["nu"]), force=tuple(d["force"]), geo=geo,
                      macro_real=_abi.REAL_F64 if st == _abi.STORE_F64 else _abi.REAL_F32, threads=4,
                      store_range=1.0, store_range_class=tuple(d["store_range_class"]))
    h.set_inflow(*d["inflow"])
    h.import_f64(d["f0"], which=0)
    ha = None
    if "ade_geo" in d:
        ha = O.HostAde(h, collision=_abi.ADE_COLL_CLBM, diffusivity=float(d["ade_D"]), geo=d["ade_geo"])
        ha.set_equilibrium(d["ade_phi0"])
        ha.set_source(d["ade_source"])
    steps = int(d["steps"])
    for t in range(steps):
        if ha is not None:
            h.step_coupled(ha, output_macro=t == steps - 1)
        else:
            h.step(output_macro=t == steps - 1)
    out = {"fT": h.export_f64(), "macros": np.stack(h.macros()).astype(np.float64)}
    if ha is not None:
        out["ade_gT"] = ha.export_f64()
    return out


def test_fixtures_present():
    assert len(FILES) >= 12


@pytest.mark.parametrize("path", FILES, ids=[p.stem for p in FILES])
def test_oracle_reproduces_fixture(path):
    d = dict(np.load(path))
    out = run_oracle(d)
    for k, v in out.items():
        assert np.array_equal(v, d[k]), (path.stem, k, np.abs(v - d[k]).max())
