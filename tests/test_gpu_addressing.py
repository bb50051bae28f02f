"""The 32-bit slot-offset form of the map, coupled, recording and FP64
adjoint kernels (a per-direction base plus a 32-bit offset, taken where a
population array holds at most 4 GiB) against the 64-bit form of the same
kernels, forced by tnl_lbm_set_wide_addressing: the same bits.  The
lattices step every row through the kernel with all boundary codes
(row_lists=False), the kernels these forms were measured on."""
import ctypes as C

import numpy as np
import pytest
import torch

from pytnl_lbm import ADE_GEO, D3Q7Lattice, D3Q27Lattice, _abi
from pytnl_lbm._abi import GEO
from pytnl_lbm.lattice import _ptr, _stream_ptr

pytestmark = pytest.mark.gpu


def duct(X, Y, Z):
    geo = np.full((X, Y, Z), GEO.FLUID, dtype=np.int16)
    geo[:, 0, :] = geo[:, -1, :] = geo[:, :, 0] = geo[:, :, -1] = GEO.WALL
    geo[0, 1:-1, 1:-1] = GEO.INFLOW
    geo[-1, 1:-1, 1:-1] = GEO.OUTFLOW_EQ
    return geo


def fields(X, Y, Z, seed):
    rng = np.random.default_rng(seed)
    return [1 + 0.01 * rng.standard_normal((X, Y, Z))] + list(0.03 * rng.standard_normal((3, X, Y, Z)))


def both_ways(run):
    """run() under the default addressing, then under the forced 64-bit form."""
    lib = _abi.load_library()
    prev = lib.tnl_lbm_set_wide_addressing(0)
    try:
        narrow = run()
        assert lib.tnl_lbm_set_wide_addressing(1) == 0
        wide = run()
    finally:
        lib.tnl_lbm_set_wide_addressing(prev)
    return narrow, wide


def bits(t):
    return t.view(torch.int16) if t.element_size() == 2 else t


@pytest.mark.parametrize("storage", ["f16dev", "bf16dev", "i16dev"])
@pytest.mark.parametrize("pattern", ["AA", "AB"])
def test_map_kernel_two_byte_forms(storage, pattern):
    X, Y, Z = 6, 70, 5

    def run():
        g = D3Q27Lattice(X, Y, Z, storage=storage, pattern=pattern, nu=0.02, geo=duct(X, Y, Z), store_range=0.5,
                         row_lists=False)
        g.set_inflow(0.03, 0.0, 0.0)
        g.set_equilibrium_field(*fields(X, Y, Z, 3))
        for t in range(4):
            g.step(output_macro=(t == 3))
        torch.cuda.synchronize()
        return [d.clone() for d in g.df] + [g.macro.clone()]

    for a, b in zip(*both_ways(run)):
        assert torch.equal(bits(a), bits(b))


@pytest.mark.parametrize("storage", ["f16dev", "i16dev"])
@pytest.mark.parametrize("pattern", ["AA", "AB"])
def test_coupled_kernel_two_byte_forms(storage, pattern):
    X, Y, Z = 6, 70, 5
    ageo = np.full((X, Y, Z), ADE_GEO.PERIODIC, dtype=np.int16)
    geo = np.full((X, Y, Z), GEO.PERIODIC, dtype=np.int16)
    geo[:, :, 0] = geo[:, :, -1] = GEO.WALL
    ageo[:, :, 0] = ageo[:, :, -1] = ADE_GEO.WALL

    def run():
        g = D3Q27Lattice(X, Y, Z, storage=storage, pattern=pattern, nu=0.02, geo=geo, store_range=0.5, row_lists=False)
        a = D3Q7Lattice(g, collision="CLBM", diffusivity=0.0125, geo=ageo, phi_ref=0.0, store_scale=1000.0)
        g.set_equilibrium_field(*fields(X, Y, Z, 4))
        a.set_equilibrium(torch.tensor(np.random.default_rng(5).random((X, Y, Z)) * 1e-3))
        for t in range(4):
            g.step_coupled(a, output_macro=(t == 3))
        torch.cuda.synchronize()
        return [d.clone() for d in g.df] + [d.clone() for d in a.df] + [g.macro.clone()]

    for a, b in zip(*both_ways(run)):
        assert torch.equal(bits(a), bits(b))


@pytest.mark.parametrize("storage,tape_f64", [("f64", 1), ("f64", 0), ("f32", 0)])
def test_recording_step(storage, tape_f64):
    from pytnl_lbm.adjoint import tape_elements
    X, Y, Z = 6, 70, 5

    def run():
        g = D3Q27Lattice(X, Y, Z, storage=storage, pattern="AB", nu=0.02, geo=duct(X, Y, Z), row_lists=False)
        g.set_inflow(0.03, 0.0, 0.0)
        g.set_equilibrium_field(*fields(X, Y, Z, 6))
        tape = torch.zeros(tape_elements(g.d.XYZ), dtype=torch.float64 if tape_f64 else torch.float32, device="cuda")
        g.d.even_iter = 1
        g._check(g.lib.tnl_lbm_d3q27_step_record(C.byref(g.cfg), C.byref(g.d), _ptr(tape), tape_f64, _stream_ptr()))
        torch.cuda.synchronize()
        return [g.df[1].clone(), tape]

    for a, b in zip(*both_ways(run)):
        assert torch.equal(a, b)


@pytest.mark.parametrize("case", ["periodic", "duct"])
def test_fp64_adjoint(case):
    """Periodic: every slot is written by plain stores, bit for bit.  Duct:
    the face fix-up adds with atomics whose order varies from run to run, so
    the two forms agree to FP64 rounding there."""
    X, Y, Z = 6, 70, 5

    def run():
        g = D3Q27Lattice(X, Y, Z, pattern="AB", nu=0.02, geo=None if case == "periodic" else duct(X, Y, Z))
        g.set_inflow(0.03, 0.0, 0.0)
        g.set_equilibrium_field(*fields(X, Y, Z, 8))
        gen = torch.Generator(device="cuda").manual_seed(9)
        lam_next = torch.randn(g.df[0].numel(), generator=gen, device="cuda", dtype=torch.float64)
        lam = torch.zeros_like(lam_next)
        grad = torch.zeros(3 * Y * Z, dtype=torch.float64, device="cuda")
        g.adjoint_step(lam_next, lam, grad, which=0)
        torch.cuda.synchronize()
        return [lam, grad]

    narrow, wide = both_ways(run)
    for a, b in zip(narrow, wide):
        if case == "periodic":
            assert torch.equal(a, b)
        else:
            assert (a - b).abs().max().item() <= 1e-13 * max(b.abs().max().item(), 1e-300)
