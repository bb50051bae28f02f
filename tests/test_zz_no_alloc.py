"""The allocation-free reduction (tnl_lbm_d3q27_reduce_macro_noalloc) through
the C ABI: the same six FP64 sums as tnl_lbm_d3q27_reduce_macro, bit for bit,
on two calls in a row, the second on a smaller lattice (its file-scope scratch
keeps the first call's partial sums, which must not leak into the second)."""
import ctypes as C

import numpy as np
import pytest
import torch

from pytnl_lbm import D3Q27Lattice, _abi, tgv_fields

pytestmark = pytest.mark.gpu


def noalloc(lat):
    lib = _abi.load_library()
    fn = lib.tnl_lbm_d3q27_reduce_macro_noalloc
    fn.argtypes = [C.POINTER(_abi.Config), C.POINTER(_abi.D3Q27Data), C.c_void_p, C.c_void_p]
    fn.restype = C.c_int
    out = torch.full((6,), float("nan"), dtype=torch.float64, device=lat.device)
    stream = torch.cuda.current_stream(lat.device).cuda_stream
    rc = fn(C.byref(lat.cfg), C.byref(lat.d), C.c_void_p(out.data_ptr()), C.c_void_p(stream))
    _abi.check(lib, rc)
    torch.cuda.synchronize()
    return out.cpu().numpy()


def run(shape, storage, steps, seed):
    X, Y, Z = shape
    mreal = _abi.REAL_F64 if storage == "f64" else _abi.REAL_F32
    lat = D3Q27Lattice(X, Y, Z, storage=storage, nu=0.01, macro_real=mreal, device="cuda:0")
    rho, vx, vy, vz = tgv_fields(X, Y, Z, 0.05)
    rng = np.random.default_rng(seed)
    lat.set_equilibrium_field(rho * (1 + 0.01 * rng.standard_normal(rho.shape)), vx, vy, vz)
    for t in range(steps):
        lat.step(output_macro=(t == steps - 1))
    torch.cuda.synchronize()
    print(f"no-alloc reduction: {shape} {storage}", flush=True)
    ref = np.array(lat.reduce_macro())
    got = noalloc(lat)
    return ref, got


@pytest.mark.parametrize("storage", ["f64", "f32"])
def test_noalloc_reduction_matches_reduce_macro_on_two_calls(storage):
    ref1, got1 = run((40, 48, 36), storage, 3, 1)
    ref2, got2 = run((6, 10, 4), storage, 2, 2)
    assert np.array_equal(ref1, got1), (ref1, got1)
    assert np.array_equal(ref2, got2), (ref2, got2)
    assert ref2[5] == 6 * 10 * 4 and ref1[5] == 40 * 48 * 36
    assert not np.array_equal(got1, got2)
