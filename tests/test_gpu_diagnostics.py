"""GPU parity of the FP64 diagnostics of sim_4.cu:257-300 beyond mass and
energy: the velocity-gradient reductions (tnl_lbm_d3q27_reduce_gradients),
sum rho |G|^2 with computeGradVel's central differences (sim_4.cu:20-60) and
sum rho |curl u|^2 (computeVorticity, sim_4.cu:62-82).

Pinned three ways: a closed form on the discrete Taylor-Green field (the
central difference of sin(k i) is sin(k) cos(k i), and every product of three
squared sines or cosines averages to 1/8 over a period of at least 3 sites);
numpy's central differences on the oracle's own macroscopic fields after
several steps; and a distributed slab whose overlap planes hold the periodic
neighbours, which must give the periodic box's sums."""
import numpy as np
import pytest
import torch

import oracle as O
from lbm_common import rel_max, stream_periodic
from pytnl_lbm import D3Q27Lattice, _abi, tgv_fields

pytestmark = pytest.mark.gpu


def numpy_gradients(rho, u):
    """sum rho |G|^2, sum rho |curl u|^2 with periodic central differences."""
    G = np.empty((3, 3) + rho.shape)
    for a in range(3):
        for b in range(3):
            G[a, b] = 0.5 * (np.roll(u[a], -1, axis=b) - np.roll(u[a], 1, axis=b))
    w = np.stack([G[2, 1] - G[1, 2], G[0, 2] - G[2, 0], G[1, 0] - G[0, 1]])
    return float((rho * (G ** 2).sum(axis=(0, 1))).sum()), float((rho * (w ** 2).sum(axis=0)).sum())


@pytest.mark.parametrize("shape", [(16, 16, 16), (12, 20, 9)])
@pytest.mark.parametrize("storage", ["f64", "f32"])
def test_gradient_sums_of_the_discrete_taylor_green_field(shape, storage):
    X, Y, Z = shape
    U0 = 0.05
    lat = D3Q27Lattice(X, Y, Z, storage=storage, nu=0.01)
    _, vx, vy, vz = tgv_fields(X, Y, Z, U0)
    lat.set_equilibrium_field(np.ones((X, Y, Z)), vx, vy, vz)
    lat.compute_macro()
    g2, w2 = lat.reduce_gradients()
    sx, sy, sz = (np.sin(2 * np.pi / n) for n in shape)
    N = X * Y * Z
    g_exact = U0 ** 2 * N / 8 * 2 * (sx ** 2 + sy ** 2 + sz ** 2)
    w_exact = U0 ** 2 * N / 8 * (2 * sz ** 2 + (sx + sy) ** 2)
    tol = 1e-12 if storage == "f64" else 1e-6
    assert abs(g2 / g_exact - 1) < tol, (g2, g_exact)
    assert abs(w2 / w_exact - 1) < tol, (w2, w_exact)


def test_gradient_sums_match_numpy_on_the_oracle_fields():
    """A seeded random flow stepped 12 times by the GPU and the oracle side
    by side; the GPU's sums against numpy's on the oracle's macros."""
    X, Y, Z = 10, 14, 8
    nu = 0.01
    rng = np.random.default_rng(21)
    r = (1 + 0.01 * rng.standard_normal((X, Y, Z)),) + tuple(0.03 * rng.standard_normal((3, X, Y, Z)))
    ref = O.HostLattice(X, Y, Z, pattern=_abi.PATTERN_AB, nu=nu)
    ref.set_equilibrium_field(*r)
    gpu = D3Q27Lattice(X, Y, Z, nu=nu)
    gpu.import_f64(stream_periodic(ref.export_f64(0)))
    for t in range(12):
        gpu.step(output_macro=True)
        ref.step(output_macro=True)
        if t % 4 == 3:
            m = ref.macros()
            g_ref, w_ref = numpy_gradients(m[0], np.stack(m[1:]))
            g2, w2 = gpu.reduce_gradients()
            assert abs(g2 / g_ref - 1) < 1e-11 and abs(w2 / w_ref - 1) < 1e-11, (t, g2, g_ref, w2, w_ref)


def test_distributed_slab_reads_its_overlap_planes():
    """With TNL_LBM_FLAG_DISTRIBUTED_X the x-neighbours of the first and last
    planes come from the overlap planes: filled with the periodic neighbours,
    the slab's sums equal the periodic box's."""
    X, Y, Z = 9, 11, 7
    rng = np.random.default_rng(4)
    rho = 1 + 0.01 * rng.standard_normal((X, Y, Z))
    u = 0.03 * rng.standard_normal((3, X, Y, Z))
    box = D3Q27Lattice(X, Y, Z, nu=0.01)
    slab = D3Q27Lattice(X, Y, Z, nu=0.01, overlap=(1, 0, 0))
    for lat, ghost in ((box, False), (slab, True)):
        fields = [rho, *u]
        if ghost:  # storage x = 0 and X + 1 hold planes X - 1 and 0
            fields = [np.concatenate([a[-1:], a, a[:1]], axis=0) for a in fields]
        m = lat.macro.view(4, lat.d.XYZ)
        Xs = X + (2 if ghost else 0)
        for k, a in enumerate(fields):
            m[k, :Xs * Y * Z] = torch.from_numpy(np.ascontiguousarray(a.transpose(0, 2, 1)).reshape(-1)).to(m)
    g_box = box.reduce_gradients()
    g_slab = slab.reduce_gradients(flags=_abi.FLAG_DISTRIBUTED_X)
    g_np = numpy_gradients(rho, u)
    assert rel_max(np.array(g_box), np.array(g_np)) < 1e-13
    assert rel_max(np.array(g_slab), np.array(g_box)) < 1e-15
    # without the flag the slab wraps inside its interior planes instead
    assert rel_max(np.array(slab.reduce_gradients()), np.array(g_box)) < 1e-15
