"""The HIP path against the committed whole-step fixtures of tests/golden/
(oracle outputs from the same initial populations): FP64 to 1e-12 of the
field scale, FP32 deviation storage to 5e-6, binary16 to 3e-3, and the
16-bit fixed point to a few quanta of its class bounds."""
from pathlib import Path

import numpy as np
import pytest
import torch

from lbm_common import CX, CY, CZ, W, rel_max
from pytnl_lbm import D3Q7Lattice, D3Q27Lattice, _abi

pytestmark = pytest.mark.gpu

GOLD = Path(__file__).resolve().parent / "golden"
FILES = sorted(GOLD.glob("*.npz"))
NAMES = {_abi.STORE_F64: "f64", _abi.STORE_F32_DEV: "f32dev", _abi.STORE_F16_DEV: "f16dev", _abi.STORE_I16_DEV: "i16dev"}


@pytest.mark.parametrize("path", FILES, ids=[p.stem for p in FILES])
def test_gpu_matches_fixture(path):
    d = dict(np.load(path))
    X, Y, Z = d["f0"].shape[1:]
    st = int(d["storage"])
    g = D3Q27Lattice(X, Y, Z, collision=int(d["collision"]), storage=NAMES[st], pattern=int(d["pattern"]),
                     nu=float(d["nu"]), force=tuple(d["force"]), geo=d.get("geo"),
                     macro_real=_abi.REAL_F64 if st == _abi.STORE_F64 else _abi.REAL_F32,
                     store_range=1.0, store_range_class=tuple(d["store_range_class"]))
    g.set_inflow(*d["inflow"])
    g.import_f64(d["f0"], which=0)
    a = None
    if "ade_geo" in d:
        a = D3Q7Lattice(g, collision="CLBM", diffusivity=float(d["ade_D"]), geo=d["ade_geo"])
        a.set_equilibrium(d["ade_phi0"])
        a.set_source(d["ade_source"])
    steps = int(d["steps"])
    for t in range(steps):
        if a is not None:
            g.step_coupled(a, output_macro=t == steps - 1)
        else:
            g.step(output_macro=t == steps - 1)
    torch.cuda.synchronize()
    fT = g.export_f64().cpu().numpy()
    macros = np.stack([m.double().cpu().numpy() for m in g.macros()])
    if st == _abi.STORE_I16_DEV:
        cls = np.array([int(c[0] != 0) + int(c[1] != 0) + int(c[2] != 0) for c in zip(CX, CY, CZ)])
        quantum = (W * d["store_range_class"][cls] / 32767)[:, None, None, None]
        diff = np.abs(fT - d["fT"])
        assert diff.max() <= 4 * quantum.max()
        assert np.mean(diff > 0.5 * quantum) < 0.05
        return
    tol = {_abi.STORE_F64: 1e-12, _abi.STORE_F32_DEV: 5e-6, _abi.STORE_F16_DEV: 3e-3}[st]
    Wb = W[:, None, None, None]
    scale = np.abs(d["fT"] - Wb).max()  # deviations from rest set the scale of the populations' rounding
    assert np.abs(fT - d["fT"]).max() <= tol * scale, (path.stem, np.abs(fT - d["fT"]).max() / scale)
    m = d["macros"]
    for k in range(4):
        ref = m[k] - (1.0 if k == 0 else 0.0)
        assert np.abs(macros[k] - m[k]).max() <= max(tol, 1e-12) * max(np.abs(ref).max(), 1e-3), (path.stem, k)
    if a is not None:
        ga = a.export_f64().cpu().numpy()
        assert rel_max(ga, d["ade_gT"]) < tol, rel_max(ga, d["ade_gT"])
