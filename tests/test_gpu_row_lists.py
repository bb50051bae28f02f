"""Row-list launches (tnl_lbm_classify_rows): the rows of FLUID/PERIODIC sites
(class 1) and of those with WALL sites (class 2) run through the lean
kernels, the other rows (class 0) through the kernel with every boundary
code.  The classification is checked against numpy, and every list-mode
step against the same step with all rows in the general kernel
(row_lists=False), for every storage form, both patterns, the coupled
NSE + D3Q7 step and the recording step, one step at a time from the same
state (the general lattice's populations are copied into the other before
each step).

The lean kernels run the same source arithmetic as the general one, but the
compiler fuses multiplies and adds differently in the two code shapes, so
one step agrees to the rounding of the stored type rather than bit for bit:
float arrays within a few epsilons of their type times their largest
magnitude (the macros' velocity components pass through zero, where a count
of units in the last place says nothing), the 2-byte floating-point forms'
stored values within one epsilon of their largest magnitude (one unit in
the last place at the top of the range; near zero, where the deviations
cancel, float32 rounding spans a few of binary16's small units), the
fixed-point form's stored integers within one unit."""
import ctypes as C

import numpy as np
import pytest
import torch

from pytnl_lbm import ADE_GEO, D3Q7Lattice, D3Q27Lattice, _abi, i16_class_ranges
from pytnl_lbm._abi import GEO
from pytnl_lbm.lattice import _ptr, _stream_ptr

pytestmark = pytest.mark.gpu

STORAGES = ["f64", "f32", "f32dev", "f16dev", "bf16dev", "i16dev"]
SHAPES = [(6, 70, 5), (4, 128, 6)]  # linear mapping over the lists; whole rows per block


def mixed_geometry(X, Y, Z):
    """Periodic box, walls on the z faces (class 2 rows), a wall obstacle and
    some FLUID sites in the interior, and rows holding other codes (class 0)."""
    geo = np.full((X, Y, Z), GEO.PERIODIC, dtype=np.int16)
    geo[:, :, 0] = geo[:, :, -1] = GEO.WALL
    geo[1, Y // 3:Y // 2, 2] = GEO.WALL           # obstacle: rows of PERIODIC and WALL
    geo[2, :Y // 4, 1:Z - 1] = GEO.FLUID           # FLUID beside PERIODIC in one row (class 1)
    geo[X - 1, 5, 1] = GEO.SYM_TOP                 # one other code: its row is class 0
    geo[0, 3:9, 2] = GEO.OUTFLOW_EQ
    return geo


def ade_geometry(X, Y, Z):
    geo = np.full((X, Y, Z), ADE_GEO.PERIODIC, dtype=np.int16)
    geo[:, :, 0] = geo[:, :, -1] = ADE_GEO.WALL
    geo[2, Y // 2, 1:Z - 1] = ADE_GEO.WALL
    geo[X - 2, 7, 3] = ADE_GEO.SYM_BOTTOM
    return geo


def expected_lists(X, Z, nse_geo, ade_geo=None):
    def cls(codes, fluid, periodic, wall):
        if np.all((codes == fluid) | (codes == periodic)):
            return 1
        if np.all((codes == fluid) | (codes == periodic) | (codes == wall)):
            return 2
        return 0
    lists = {0: [], 1: [], 2: []}
    for x in range(X):
        for z in range(Z):
            c = cls(nse_geo[x, :, z], GEO.FLUID, GEO.PERIODIC, GEO.WALL)
            if ade_geo is not None:
                a = cls(ade_geo[x, :, z], ADE_GEO.FLUID, ADE_GEO.PERIODIC, ADE_GEO.WALL)
                c = 0 if (c == 0 or a == 0) else max(c, a)
            lists[c].append((x << 16) | z)
    return lists


def got_lists(rows, counts):
    r = rows.cpu().numpy().astype(np.int64) & 0xFFFFFFFF
    out, off = {}, 0
    for c in range(3):
        out[c] = list(r[off:off + counts[c]])
        off += counts[c]
    return out


def units(a: torch.Tensor, b: torch.Tensor) -> float:
    """Largest difference of two arrays of one element type in units of the
    type's epsilon times the array's largest magnitude (fixed point: the
    plain difference of the stored integers)."""
    if not a.dtype.is_floating_point:
        return float((a.to(torch.int64) - b.to(torch.int64)).abs().max())
    eps = torch.finfo(a.dtype).eps
    a, b = a.double(), b.double()
    scale = max(float(a.abs().max()), 1e-300)
    return float((a - b).abs().max()) / (eps * scale)


def within(a, b):
    """One step from the same state: a few epsilons for the FP64 and FP32
    arrays, one for the stored 2-byte values."""
    return units(a, b) <= (1.0 if a.element_size() == 2 else 16.0)


def fields(X, Y, Z, seed):
    rng = np.random.default_rng(seed)
    return [1 + 0.01 * rng.standard_normal((X, Y, Z))] + list(0.03 * rng.standard_normal((3, X, Y, Z)))


def lattice(X, Y, Z, storage, pattern, geo, row_lists, **kw):
    rc = i16_class_ranges(0.2) if storage == "i16dev" else None
    mreal = _abi.REAL_F64 if storage == "f64" else _abi.REAL_F32
    return D3Q27Lattice(X, Y, Z, storage=storage, pattern=pattern, nu=0.01, geo=geo, row_lists=row_lists,
                        macro_real=mreal, store_range_class=rc, **kw)


@pytest.mark.parametrize("shape", SHAPES)
def test_classification_matches_numpy(shape):
    X, Y, Z = shape
    geo = mixed_geometry(X, Y, Z)
    lat = lattice(X, Y, Z, "f64", "AA", geo, True)
    counts = tuple(lat.d.row_count)
    exp = expected_lists(X, Z, geo)
    assert counts == (len(exp[0]), len(exp[1]), len(exp[2]))
    assert all(counts), counts  # the geometry holds every class
    assert got_lists(lat.rows, counts) == exp
    ade = D3Q7Lattice(lat, collision="CLBM", diffusivity=0.02, geo=ade_geometry(X, Y, Z))
    ace = tuple(ade.d.row_count)
    exp2 = expected_lists(X, Z, geo, ade_geometry(X, Y, Z))
    assert ace == (len(exp2[0]), len(exp2[1]), len(exp2[2]))
    assert got_lists(ade.rows, ace) == exp2
    # a changed map is classified again
    geo2 = geo.copy()
    geo2[:, :, 2] = GEO.INFLOW
    lat.set_geometry(geo2)
    exp3 = expected_lists(X, Z, geo2)
    assert got_lists(lat.rows, tuple(lat.d.row_count)) == exp3
    # ... also after an in-place change and refresh_geometry
    lat.map.zero_()  # all FLUID
    lat.refresh_geometry()
    assert tuple(lat.d.row_count) == (0, X * Z, 0)


@pytest.mark.parametrize("pattern", ["AA", "AB"])
@pytest.mark.parametrize("storage", STORAGES)
@pytest.mark.parametrize("shape", SHAPES)
def test_list_steps_match_general_kernel(shape, storage, pattern):
    X, Y, Z = shape
    geo = mixed_geometry(X, Y, Z)
    a = lattice(X, Y, Z, storage, pattern, geo, True, force=(1e-5, 0.0, -2e-6))
    b = lattice(X, Y, Z, storage, pattern, geo, False, force=(1e-5, 0.0, -2e-6))
    assert all(a.d.row_count) and not b.d.rows
    r = fields(X, Y, Z, 3)
    a.set_equilibrium_field(*r)
    b.set_equilibrium_field(*r)
    worst = 0
    for t in range(6):
        for x, y in zip(a.df, b.df):
            x.copy_(y)
        a.step(output_macro=True)
        b.step(output_macro=True)
        torch.cuda.synchronize()
        for k in range(len(a.df)):
            assert within(a.df[k], b.df[k]), (storage, pattern, shape, t, k, units(a.df[k], b.df[k]))
            worst = max(worst, units(a.df[k], b.df[k]))
        assert within(a.macro, b.macro), (storage, pattern, shape, t, units(a.macro, b.macro))
    print(f"{storage} {pattern} {shape}: populations within {worst:.2f} units")


@pytest.mark.parametrize("pattern", ["AA", "AB"])
@pytest.mark.parametrize("storage", ["f64", "f32", "f16dev", "i16dev"])
def test_coupled_list_steps_match_general_kernel(storage, pattern):
    X, Y, Z = 5, 70, 6
    geo, ageo = mixed_geometry(X, Y, Z), ade_geometry(X, Y, Z)
    pair = []
    for rl in (True, False):
        nse = lattice(X, Y, Z, storage, pattern, geo, rl)
        ade = D3Q7Lattice(nse, collision="CLBM", diffusivity=0.02, geo=ageo, phi_ref=0.1, store_scale=50.0)
        nse.set_equilibrium_field(*fields(X, Y, Z, 5))
        ade.set_equilibrium(0.1)
        src = np.random.default_rng(7).standard_normal((X, Y, Z)) * 1e-3
        ade.set_source(src)
        pair.append((nse, ade))
    (na, aa), (nb, ab) = pair
    assert all(aa.d.row_count) and not ab.d.rows
    worst = 0
    for t in range(6):
        for x, y in list(zip(na.df, nb.df)) + list(zip(aa.df, ab.df)):
            x.copy_(y)
        na.step_coupled(aa, output_macro=True)
        nb.step_coupled(ab, output_macro=True)
        torch.cuda.synchronize()
        for x, y in list(zip(na.df, nb.df)) + list(zip(aa.df, ab.df)) + [(na.macro, nb.macro), (aa.phi, ab.phi)]:
            assert within(x, y), (storage, pattern, t, units(x, y))
            worst = max(worst, units(x, y))
    print(f"coupled {storage} {pattern}: within {worst:.2f} units")


@pytest.mark.parametrize("storage,tape_f64", [("f64", 1), ("f64", 0), ("f32", 0)])
def test_recording_step_ignores_the_lists(storage, tape_f64):
    X, Y, Z = 6, 70, 5
    geo = mixed_geometry(X, Y, Z)
    geo[0, 3:9, 2] = GEO.PERIODIC  # the recording step's codes: no equilibrium outflow needed here
    out = []
    for rl in (True, False):
        lat = lattice(X, Y, Z, storage, "AB", geo, rl)
        lat.set_equilibrium_field(*fields(X, Y, Z, 9))
        n = (lat.d.XYZ + 63) // 64 * 640
        tape = torch.zeros(n, dtype=torch.float64 if tape_f64 else torch.float32, device="cuda")
        lat.update_kernel_data()
        lat._check(lat.lib.tnl_lbm_d3q27_step_record(C.byref(lat.cfg), C.byref(lat.d), _ptr(tape), tape_f64,
                                                     _stream_ptr()))
        torch.cuda.synchronize()
        out.append((lat.df[1].clone(), tape))
    assert within(out[0][0], out[1][0]), units(out[0][0], out[1][0])
    assert within(out[0][1], out[1][1]), units(out[0][1], out[1][1])


def test_sub_box_launch_ignores_the_lists():
    """A launch over part of the box steps that part only, lists or not."""
    X, Y, Z = 6, 70, 5
    geo = mixed_geometry(X, Y, Z)
    a = lattice(X, Y, Z, "f64", "AB", geo, True)
    b = lattice(X, Y, Z, "f64", "AB", geo, False)
    r = fields(X, Y, Z, 4)
    a.set_equilibrium_field(*r)
    b.set_equilibrium_field(*r)
    for lat in (a, b):
        lat.df[1].zero_()
        lat.step(offset=(1, 0, 1), end=(X - 1, Y, Z - 1))
    torch.cuda.synchronize()
    assert torch.equal(a.df[1], b.df[1])


@pytest.mark.parametrize("pattern", ["AA", "AB"])
@pytest.mark.parametrize("storage", ["f16dev", "i16dev", "f64"])
def test_slab_with_overlaps(storage, pattern):
    """An x-slab with one ghost plane per side and a map (walls on the z faces,
    an obstacle), stepped with DISTRIBUTED_X: the rows are classified over
    the local sites and the lean kernels read their x neighbours from the
    ghost planes as the general kernel does.  Both lattices start from the
    same stored arrays, ghosts included."""
    X, Y, Z = 4, 70, 6
    geo = mixed_geometry(X, Y, Z)
    a = lattice(X, Y, Z, storage, pattern, geo, True, overlap=(1, 0, 0))
    b = lattice(X, Y, Z, storage, pattern, geo, False, overlap=(1, 0, 0))
    assert a.map.numel() == (X + 2) * Y * Z
    assert all(a.d.row_count) and sum(a.d.row_count) == X * Z
    r = fields(X + 2, Y, Z, 6)
    whole = D3Q27Lattice(X + 2, Y, Z, storage=storage, pattern=pattern, nu=0.01,
                         store_range_class=i16_class_ranges(0.2) if storage == "i16dev" else None)
    whole.set_equilibrium_field(*r)  # the stored arrays of a (X+2, Y, Z) box: interior and ghosts
    for t in range(4):
        for x, y, w in zip(a.df, b.df, whole.df):
            x.copy_(w)
            y.copy_(w)
        a.step(output_macro=True, flags=_abi.FLAG_DISTRIBUTED_X)
        b.step(output_macro=True, flags=_abi.FLAG_DISTRIBUTED_X)
        torch.cuda.synchronize()
        for x, y in zip(a.df, b.df):
            assert within(x, y), (storage, pattern, t, units(x, y))
        assert within(a.macro, b.macro), (storage, pattern, t, units(a.macro, b.macro))
        whole.step()
