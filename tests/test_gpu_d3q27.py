"""GPU parity: libtnl_lbm_amd.so (gfx950) against the CPU oracle on the same
seeded inputs.  FP64 is compared per step to ~1e-13 (the GPU contracts
multiply-adds into FMAs, the oracle does not); the laminar Taylor-Green
vortex of sim_4.cu must match the reference's FP64 A-B path to <= 1e-10
max relative error in rho and u over 1000 steps."""
import numpy as np
import pytest
import torch

import oracle as O
from lbm_common import CX, CY, CZ, OPP, W, rel_max, stream_periodic
from pytnl_lbm import D3Q27Lattice, _abi, tgv_fields
from pytnl_lbm._abi import GEO

pytestmark = pytest.mark.gpu

STORE = {"f64": _abi.STORE_F64, "f32": _abi.STORE_F32, "f32dev": _abi.STORE_F32_DEV,
         "f16dev": _abi.STORE_F16_DEV, "bf16dev": _abi.STORE_BF16_DEV, "i16dev": _abi.STORE_I16_DEV}
PAT = {"AA": _abi.PATTERN_AA, "AB": _abi.PATTERN_AB}


def fields(X, Y, Z, seed, amp=0.03):
    rng = np.random.default_rng(seed)
    rho = 1 + 0.01 * rng.standard_normal((X, Y, Z))
    u = amp * rng.standard_normal((3, X, Y, Z))
    return rho, u[0], u[1], u[2]


def pair(X, Y, Z, coll=0, storage="f64", pattern="AA", geo=None, nu=0.01, force=(0.0, 0.0, 0.0), store_range=1.0,
         store_range_class=None):
    mreal = _abi.REAL_F64 if storage == "f64" else _abi.REAL_F32
    g = D3Q27Lattice(X, Y, Z, collision=coll, storage=storage, pattern=pattern, nu=nu, force=force, geo=geo,
                     macro_real=mreal, store_range=store_range, store_range_class=store_range_class)
    h = O.HostLattice(X, Y, Z, collision=coll, storage=STORE[storage], pattern=PAT[pattern], nu=nu, force=force,
                      geo=geo, macro_real=mreal, store_range=store_range, store_range_class=store_range_class)
    return g, h


def cmp_pops(g, h):
    return rel_max(g.export_f64().cpu().numpy(), h.export_f64())


@pytest.mark.parametrize("coll", [0, 1, 2])
@pytest.mark.parametrize("shape", [(6, 7, 5), (3, 70, 4), (2, 130, 3)])
def test_periodic_aa_f64_matches_oracle(coll, shape):
    X, Y, Z = shape
    g, h = pair(X, Y, Z, coll, force=(1e-5, -2e-5, 3e-6))
    r = fields(X, Y, Z, 1)
    g.set_equilibrium_field(*r)
    h.set_equilibrium_field(*r)
    assert cmp_pops(g, h) < 1e-15  # FMA-contracted equilibrium vs the uncontracted oracle
    for t in range(8):
        g.step(output_macro=True)
        h.step(output_macro=True)
        torch.cuda.synchronize()
        assert cmp_pops(g, h) < 1e-13, (coll, shape, t)
        for a, b in zip(g.macros(), h.macros()):
            assert rel_max(a.cpu().numpy(), b) < 1e-13


@pytest.mark.parametrize("pattern", ["AA", "AB"])
def test_general_path_periodic_map_matches(pattern):
    """All-PERIODIC map through the general (map-reading) kernel."""
    X, Y, Z = 5, 9, 4
    geo = np.full((X, Y, Z), GEO.PERIODIC, dtype=np.int16)
    g, h = pair(X, Y, Z, 0, pattern=pattern, geo=geo)
    r = fields(X, Y, Z, 2)
    g.set_equilibrium_field(*r)
    h.set_equilibrium_field(*r)
    for t in range(6):
        g.step(output_macro=True)
        h.step(output_macro=True)
        assert cmp_pops(g, h) < 1e-13


def channel_geo(X, Y, Z, inflow=GEO.INFLOW, outflow=GEO.OUTFLOW_EQ):
    geo = np.full((X, Y, Z), GEO.FLUID, dtype=np.int16)
    geo[:, 1, :] = GEO.WALL
    geo[:, -2, :] = GEO.WALL
    geo[:, :, 1] = GEO.WALL
    geo[:, :, -2] = GEO.WALL
    geo[1, 2:-2, 2:-2] = inflow
    geo[-2, 2:-2, 2:-2] = outflow
    geo[0] = GEO.NOTHING
    geo[-1] = GEO.NOTHING
    geo[:, 0] = GEO.NOTHING
    geo[:, -1] = GEO.NOTHING
    geo[:, :, 0] = GEO.NOTHING
    geo[:, :, -1] = GEO.NOTHING
    return geo


@pytest.mark.parametrize("pattern,inflow,outflow", [
    ("AA", GEO.INFLOW, GEO.OUTFLOW_EQ),
    ("AA", GEO.INFLOW_EQ_LEFT, GEO.OUTFLOW_EQ),
    ("AA", GEO.INFLOW_LEFT, GEO.OUTFLOW_EQ),
    ("AB", GEO.INFLOW, GEO.OUTFLOW_RIGHT),
    ("AB", GEO.INFLOW_BB_LEFT, GEO.OUTFLOW_RIGHT_INTERP),
])
def test_channel_bcs_match_oracle(pattern, inflow, outflow):
    X, Y, Z = 12, 9, 8
    geo = channel_geo(X, Y, Z, inflow, outflow)
    g, h = pair(X, Y, Z, 0, pattern=pattern, geo=geo, nu=0.02)
    prof = np.linspace(0.01, 0.03, Y * Z).reshape(Z, Y)
    for lat in (g, h):
        lat.set_equilibrium(1.0, 0.02, 0.0, 0.0)
        lat.set_inflow(0.02, 0.001, 0.0, profile=(prof, None, None))
    for t in range(10):
        g.step(output_macro=True)
        h.step(output_macro=True)
        assert cmp_pops(g, h) < 1e-13, t
        for a, b in zip(g.macros(), h.macros()):
            assert rel_max(a.cpu().numpy(), b) < 5e-13  # small transverse velocities: FMA rounding


@pytest.mark.parametrize("pattern", ["AA", "AB"])
def test_f32_channel_with_float_inflow_profiles_matches_oracle(pattern):
    """TraitsSP's NSE_Data_InflowProfile holds float arrays (lbm_data.h:100-104):
    inflow_profile_real = F32 reads them as such; all three components vary."""
    X, Y, Z = 12, 9, 8
    geo = channel_geo(X, Y, Z, GEO.INFLOW, GEO.OUTFLOW_EQ)
    g, h = pair(X, Y, Z, 0, storage="f32", pattern=pattern, geo=geo, nu=0.02)
    rng = np.random.default_rng(5)
    prof = [(0.02 + 0.005 * rng.standard_normal((Z, Y))).astype(np.float32),
            (0.002 * rng.standard_normal((Z, Y))).astype(np.float32),
            (0.002 * rng.standard_normal((Z, Y))).astype(np.float32)]
    for lat in (g, h):
        lat.set_equilibrium(1.0, 0.02, 0.0, 0.0)
        lat.set_inflow(0.0, 0.0, 0.0, profile=prof, profile_real="f32")
    assert g.d.inflow_profile_real == _abi.REAL_F32
    for t in range(6):
        g.step(output_macro=True)
        h.step(output_macro=True)
        assert cmp_pops(g, h) < 2e-6, t


@pytest.mark.parametrize("pattern,inflow,outflow", [("AA", GEO.INFLOW, GEO.OUTFLOW_EQ),
                                                     ("AB", GEO.INFLOW_EQ_LEFT, GEO.OUTFLOW_RIGHT)])
@pytest.mark.parametrize("storage,tol", [("f32dev", 2e-5), ("f16dev", 3e-3), ("bf16dev", 2e-2)])
def test_deviation_storage_channel_bcs_match_oracle(storage, tol, pattern, inflow, outflow):
    """The deviation storage forms through the map kernel: walls, inflow,
    outflow and NOTHING sites next to fluid sites (the FP32 raw-moment
    collision at fluid sites, the reference chain at the boundary codes),
    against the oracle's FP32 deviation chain.  The bound for F32_DEV is
    four times the periodic box's: INFLOW_EQ_LEFT extrapolates
    rho = (1 + S_0 + 2 S_-x) / (1 - u_x) from O(1) FP32 quantities, whose
    rounding (contracted on the GPU, not in the oracle) is ~1e-7 absolute,
    up to ~7e-6 of this nearly uniform flow's deviation scale."""
    X, Y, Z = 12, 9, 8
    geo = channel_geo(X, Y, Z, inflow, outflow)
    g, h = pair(X, Y, Z, 0, storage=storage, pattern=pattern, geo=geo, nu=0.02)
    prof = np.linspace(0.01, 0.03, Y * Z).reshape(Z, Y)
    for lat in (g, h):
        lat.set_equilibrium(1.0, 0.02, 0.0, 0.0)
        lat.set_inflow(0.02, 0.001, 0.0, profile=(prof, None, None))
    Wb = W[:, None, None, None]
    for t in range(6):
        g.step(output_macro=True)
        h.step(output_macro=True)
        dg, dh = g.export_f64().cpu().numpy() - Wb, h.export_f64() - Wb
        assert rel_max(dg, dh) < tol, (storage, pattern, t, rel_max(dg, dh))


@pytest.mark.parametrize("sym", [GEO.SYM_TOP, GEO.SYM_BOTTOM, GEO.SYM_LEFT, GEO.SYM_RIGHT, GEO.SYM_BACK, GEO.SYM_FRONT])
def test_symmetry_planes_match_oracle(sym):
    X, Y, Z = 6, 6, 6
    geo = np.full((X, Y, Z), GEO.PERIODIC, dtype=np.int16)
    geo[2:4, 2:4, 2:4] = sym
    g, h = pair(X, Y, Z, 0, geo=geo)
    r = fields(X, Y, Z, 5)
    g.set_equilibrium_field(*r)
    h.set_equilibrium_field(*r)
    for t in range(4):
        g.step(output_macro=True)
        h.step(output_macro=True)
        assert cmp_pops(g, h) < 1e-13


def test_f32_absolute_matches_oracle():
    X, Y, Z = 6, 70, 5
    g, h = pair(X, Y, Z, 0, storage="f32")
    r = fields(X, Y, Z, 3)
    g.set_equilibrium_field(*r)
    h.set_equilibrium_field(*r)
    for t in range(6):
        g.step(output_macro=True)
        h.step(output_macro=True)
        assert cmp_pops(g, h) < 2e-6


I16_RANGE = 0.5
I16_QUANTUM = (W * I16_RANGE / 32767)[:, None, None, None]


@pytest.mark.parametrize("storage,tol", [("f32dev", 5e-6), ("f16dev", 3e-3), ("bf16dev", 2e-2), ("i16dev", None)])
def test_deviation_storage_matches_oracle(storage, tol):
    """Deviation storage against the oracle's FP32 deviation chain.  The GPU
    computes the FP32 cumulant from raw moments with FMAs; the oracle follows
    the reference's central-moment chain without contraction, so the two
    agree to FP32 rounding of the deviations (a few 1e-6 of their scale) and
    round to the same 16-bit value except near ties."""
    X, Y, Z = 6, 70, 5
    g, h = pair(X, Y, Z, 0, storage=storage, store_range=I16_RANGE)
    r = fields(X, Y, Z, 4, amp=0.02)
    g.set_equilibrium_field(*r)
    h.set_equilibrium_field(*r)
    Wb = W[:, None, None, None]
    d0 = h.export_f64() - Wb
    g0 = g.export_f64().cpu().numpy() - Wb
    if storage == "i16dev":
        # same quantised value, or the neighbour one at a rounding tie
        excess = np.abs(g0 - d0) - 1.001 * I16_QUANTUM
        assert excess.max() <= 1e-12
        assert np.mean(np.abs(g0 - d0) > 0.5 * I16_QUANTUM) < 1e-3
    else:
        assert rel_max(g0, d0) < max(tol / 4, 2.5e-7)  # <= 2 FP32 ulp
    for t in range(3):
        g.step(output_macro=True)
        h.step(output_macro=True)
        dg = g.export_f64().cpu().numpy() - Wb
        dh = h.export_f64() - Wb
        if storage == "i16dev":
            # a tie flip moves one population by one quantum; one collision
            # spreads it over the site's 27 outputs with gain ~1, and those
            # are quantised again: a few of the coarsest quanta at most
            diff = np.abs(dg - dh)
            assert diff.max() <= 4 * I16_QUANTUM.max(), (t, diff.max())
            assert np.mean(diff > 0.5 * I16_QUANTUM) < 0.02, t
        else:
            assert rel_max(dg, dh) < tol, (storage, t)


@pytest.mark.parametrize("shape", [(5, 70, 4), (3, 256, 3)])
@pytest.mark.parametrize("pattern", ["AA", "AB"])
@pytest.mark.parametrize("storage,tol", [("f16dev", 3e-3), ("bf16dev", 2e-2), ("i16dev", None)])
def test_two_byte_map_kernel_matches_oracle(storage, tol, pattern, shape):
    """The 2-byte forms through the map kernel, which addresses every slot as
    a per-direction base plus a 32-bit offset: an all-PERIODIC map (every
    face wraps) and a wall slab, with lanes laid over whole y-z planes
    (Y = 70) and over whole rows (Y = 256), against the oracle.  For I16
    every value is the oracle's or its neighbour quantum (a rounding tie
    that FP32 contraction decides the other way); each collision spreads a
    flip over the site's 27 outputs, so their share grows by ~2.5x a step
    (0.2%, 0.8%, 2%, 4.7% here; the same with 64-bit addressing)."""
    X, Y, Z = shape
    geo = np.full((X, Y, Z), GEO.PERIODIC, dtype=np.int16)
    geo[:, Y // 3, :] = GEO.WALL
    g, h = pair(X, Y, Z, 0, storage=storage, pattern=pattern, geo=geo, store_range=I16_RANGE)
    r = fields(X, Y, Z, 9, amp=0.02)
    g.set_equilibrium_field(*r)
    h.set_equilibrium_field(*r)
    Wb = W[:, None, None, None]
    for t in range(4):
        g.step(output_macro=True)
        h.step(output_macro=True)
        dg = g.export_f64().cpu().numpy() - Wb
        dh = h.export_f64() - Wb
        if storage == "i16dev":
            quanta = np.abs(dg - dh) / I16_QUANTUM
            assert quanta.max() <= 1.001, (t, quanta.max())
            assert np.mean(quanta > 0.5) < 0.01 * 3 ** t, t
        else:
            assert rel_max(dg, dh) < tol, (storage, pattern, t, rel_max(dg, dh))
        for a, b in zip(g.macros(), h.macros()):
            assert np.abs(a.cpu().numpy() - b).max() < 1e-2, t


def split_step(lat, cuts, output_macro):
    """One A-A step of the whole lattice as several launches over y ranges (the
    A-A update is in place and each slot is read and written by one site, so
    any partition gives the same result)."""
    it = lat.iterations
    for lo, hi in cuts:
        lat.iterations = it
        lat.step(output_macro=output_macro, offset=(0, lo, 0), end=(lat.X, hi, lat.Z))


@pytest.mark.parametrize("storage", ["f16dev", "bf16dev", "i16dev"])
@pytest.mark.parametrize("shape", [(6, 70, 5), (3, 260, 2), (4, 2, 3), (2, 130, 4)])
def test_paired_kernel_bitwise(storage, shape):
    """On the even A-A step the 2-byte forms update two sites per lane where
    the y range allows it (even Y, offset and length).  Launches over y
    ranges with odd bounds take the one-site kernel, so splitting the same
    step at an odd y compares the two kernels; splitting at an even y checks
    the paired kernel on partial rows.  Odd steps run the one-site kernel in
    all three lattices.
    The two kernels run the same FP32 collision, but the compiler may contract
    a different multiply-add into an FMA in each, which moves a value by one
    FP32 ulp and, at a near-tie, its 16-bit rounding by one step: allowed on
    a few slots.  A pairing mistake (halves swapped, a wrong wrap) moves whole
    populations (~1e-3) between sites."""
    X, Y, Z = shape

    def off_by(a, b):
        """Slots further apart than one 16-bit step: for the float forms one
        spacing at the larger magnitude, or binary16's subnormal spacing 2^-24
        (bfloat16: 1e-9, for values that an FP32 ulp of the populations' scale
        moves across zero)."""
        if storage == "i16dev":
            return (a.view(torch.int16).int() - b.view(torch.int16).int()).abs() > 1
        a, b = a.double(), b.double()
        step, floor = (2.0**-10, 2.0**-24) if storage == "f16dev" else (2.0**-7, 1e-9)
        return (a - b).abs() > torch.maximum(a.abs(), b.abs()) * step + floor

    lats = [D3Q27Lattice(X, Y, Z, storage=storage, nu=0.01, force=(1e-5, -2e-5, 3e-6), macro_real=_abi.REAL_F32,
                         store_range=0.5) for _ in range(3)]
    r = fields(X, Y, Z, 21, amp=0.03)
    for g in lats:
        g.set_equilibrium_field(*r)
    odd = [(0, 1), (1, Y)] if Y == 2 else [(0, Y // 2 + (Y // 2) % 2 - 1), (Y // 2 + (Y // 2) % 2 - 1, Y)]
    even = [(0, Y)] if Y == 2 else [(0, 2 * (Y // 4)), (2 * (Y // 4), Y)]
    for t in range(6):
        om = t % 2 == 1
        lats[0].step(output_macro=om)
        split_step(lats[1], odd, om)
        split_step(lats[2], even, om)
        torch.cuda.synchronize()
        for k in (1, 2):
            a, b = lats[0].df[0], lats[k].df[0]
            assert not bool(off_by(a, b).any()), (t, k)
            assert float((a.view(torch.int16) != b.view(torch.int16)).double().mean()) < 2e-3, (t, k)
            if om:
                m0, mk = lats[0].macro, lats[k].macro
                assert float((m0 - mk).abs().max()) <= 1e-6 * float(m0.abs().max()), (t, k)
        # continue both runs from the same state, so that differences do not accumulate
        for k in (1, 2):
            lats[k].df[0].copy_(lats[0].df[0])


@pytest.mark.parametrize("storage", ["f32", "f32dev", "f16dev", "i16dev"])
def test_odd_step_block_size_bitwise(storage):
    """The periodic A-A kernel of the FP32-compute forms runs 512-lane blocks
    where the row length is a multiple of 512; a launch over half the row
    takes the 256-lane blocks.  Both run the same per-site code, so stepping
    the lattice whole and as two half-row launches must give the same bits."""
    X, Y, Z = 3, 512, 2
    kw = dict(storage=storage, nu=0.01, force=(1e-5, -2e-5, 3e-6), macro_real=_abi.REAL_F32, store_range=0.5)
    a, b = D3Q27Lattice(X, Y, Z, **kw), D3Q27Lattice(X, Y, Z, **kw)
    r = fields(X, Y, Z, 41, amp=0.03)
    a.set_equilibrium_field(*r)
    b.set_equilibrium_field(*r)
    for t in range(4):
        a.step(output_macro=True)
        split_step(b, [(0, 256), (256, 512)], True)
        torch.cuda.synchronize()
        assert torch.equal(a.df[0].view(torch.int16), b.df[0].view(torch.int16)), t
        assert torch.equal(a.macro.view(torch.int32), b.macro.view(torch.int32)), t


@pytest.mark.parametrize("nslabs", [1, 2])
@pytest.mark.parametrize("pattern", ["AA", "AB"])
@pytest.mark.parametrize("storage", ["f64", "f32", "f16dev"])
def test_x_slabs_with_halo_match_the_periodic_box(storage, pattern, nslabs):
    """One periodic box stepped whole, and the same box split into x-slabs
    with a ghost plane per side, stepped with DISTRIBUTED_X and exchanged
    through tnl_lbm_d3q27_halo_x (one slab: its own periodic neighbour; two:
    a ring).  A-B pulls after a PULL exchange; A-A exchanges PULL after the
    even step and PUSH after the odd one.  Every slab must hold the whole
    box's bits after every step: the kernels read the same values from the
    ghost planes as from the wrapped neighbours, and every pushed value
    reaches its owner."""
    X, Y, Z = 8, 70, 6
    kw = dict(storage=storage, pattern=pattern, nu=0.01, force=(1e-5, -2e-5, 3e-6), macro_real=_abi.REAL_F32)
    whole = D3Q27Lattice(X, Y, Z, **kw)
    whole.set_equilibrium_field(*fields(X, Y, Z, 31, amp=0.03))
    w = X // nslabs
    slabs = [D3Q27Lattice(w, Y, Z, overlap=(1, 0, 0), **kw) for _ in range(nslabs)]
    for k, s in enumerate(slabs):
        s.stored(0)[:, 1:w + 1] = whole.stored(0)[:, k * w:(k + 1) * w]
    ring = [(slabs[k], slabs[(k + 1) % nslabs]) for k in range(nslabs)]
    for t in range(6):
        if pattern == "AB":
            for lo, up in ring:
                lo.halo_x(up, _abi.HALO_PULL)
        for s in slabs:
            s.step(flags=_abi.FLAG_DISTRIBUTED_X)
        if pattern == "AA":
            for lo, up in ring:
                lo.halo_x(up, _abi.HALO_PULL if t % 2 == 0 else _abi.HALO_PUSH)
        whole.step()
        torch.cuda.synchronize()
        ref = whole.stored(whole.current_df())
        for k, s in enumerate(slabs):
            got = s.stored(s.current_df())[:, 1:w + 1].contiguous()
            assert torch.equal(got.view(torch.uint8), ref[:, k * w:(k + 1) * w].contiguous().view(torch.uint8)), (t, k)


@pytest.mark.parametrize("pattern,geo_kind", [("AA", None), ("AB", "channel")])
def test_i16_class_ranges_match_oracle(pattern, geo_kind):
    """The fixed-point form with one bound per direction class (rest, faces,
    edges, corners): same quantised values as the oracle's store model
    (ties aside), on the periodic fast kernel and on the general kernel."""
    from pytnl_lbm import i16_class_ranges
    X, Y, Z = 6, 70, 5
    Rc = i16_class_ranges(0.08)
    cls = np.array([int(c[0] != 0) + int(c[1] != 0) + int(c[2] != 0) for c in zip(CX, CY, CZ)])
    quantum = (W * np.asarray(Rc)[cls] / 32767)[:, None, None, None]
    geo = channel_geo(X, Y, Z, outflow=GEO.OUTFLOW_RIGHT) if geo_kind else None
    g, h = pair(X, Y, Z, 0, storage="i16dev", pattern=pattern, geo=geo, store_range=0.5, store_range_class=Rc)
    r = fields(X, Y, Z, 11, amp=0.02)
    g.set_equilibrium_field(*r)
    h.set_equilibrium_field(*r)
    Wb = W[:, None, None, None]
    d0, g0 = h.export_f64() - Wb, g.export_f64().cpu().numpy() - Wb
    assert (np.abs(g0 - d0) - 1.001 * quantum).max() <= 1e-12
    for t in range(3):
        g.step(output_macro=True)
        h.step(output_macro=True)
        diff = np.abs((g.export_f64().cpu().numpy() - Wb) - (h.export_f64() - Wb))
        assert diff.max() <= 4 * quantum.max(), (t, diff.max())
        # the rest population's quantum (w_0 R_0 / 32767, ~3e-7) is only a few FP32 roundings of the
        # populations: the GPU's and the oracle's FP32 chains round to neighbouring values more often
        assert np.mean(diff > 0.5 * quantum) < 0.05, t


@pytest.mark.parametrize("storage", ["f32", "f32dev"])
def test_fp32_collision_accuracy_vs_fp64(storage):
    """The GPU's FP32 cumulant is at least as accurate as the reference's own
    chain evaluated in FP32: after 20 steps its deviation from the FP64 oracle
    is within twice the FP32 oracle's deviation from the same FP64 run."""
    X, Y, Z = 6, 70, 5
    nu = 0.01
    r = fields(X, Y, Z, 7, amp=0.02)
    g = D3Q27Lattice(X, Y, Z, storage=storage, nu=nu, macro_real=_abi.REAL_F32)
    h32 = O.HostLattice(X, Y, Z, storage=STORE[storage], nu=nu, macro_real=_abi.REAL_F32)
    h64 = O.HostLattice(X, Y, Z, nu=nu)
    for lat in (g, h32, h64):
        lat.set_equilibrium_field(*r)
    for _ in range(20):
        for lat in (g, h32, h64):
            lat.step()
    Wb = W[:, None, None, None]
    ref = h64.export_f64() - Wb
    err_gpu = rel_max(g.export_f64().cpu().numpy() - Wb, ref)
    err_cpu = rel_max(h32.export_f64() - Wb, ref)
    print(f"{storage}: GPU FP32 {err_gpu:.3e}, oracle FP32 {err_cpu:.3e} (relative to FP64)")
    assert err_gpu <= 2 * err_cpu + 1e-7


def test_reduce_macro_matches_numpy():
    X, Y, Z = 7, 65, 9
    g = D3Q27Lattice(X, Y, Z, nu=0.01)
    r = fields(X, Y, Z, 6)
    g.set_equilibrium_field(*r)
    g.compute_macro()
    out = g.reduce_macro()
    rho, vx, vy, vz = [m.cpu().numpy() for m in g.macros()]
    u2 = vx * vx + vy * vy + vz * vz
    np.testing.assert_allclose(out[0], rho.sum(), rtol=1e-13)
    np.testing.assert_allclose(out[1], (rho * u2).sum(), rtol=1e-12)
    assert out[2] == 0.0
    np.testing.assert_allclose(out[3], ((rho - 1) ** 2).sum(), rtol=1e-11)
    np.testing.assert_allclose(out[4], u2.sum(), rtol=1e-12)
    assert out[5] == X * Y * Z


def run_tgv(N, steps, every, coll, threads=16):
    """GPU A-A started from S f0 against the oracle's A-B pull scheme started
    from f0 (the reference's default pattern, defs.h:20-26)."""
    r = tgv_fields(N, N, N, 0.05)
    nu = (0.52 - 0.5) / 3.0
    h = O.HostLattice(N, N, N, collision=coll, pattern=_abi.PATTERN_AB, nu=nu, threads=threads)
    h.set_equilibrium_field(*r)
    f0 = h.export_f64(0)
    g = D3Q27Lattice(N, N, N, collision=coll, nu=nu)
    g.import_f64(stream_periodic(f0))
    worst = 0.0
    for t in range(1, steps + 1):
        out = (t % every == 0)
        g.step(output_macro=out)
        h.step(output_macro=out)
        if out:
            for a, b in zip(g.macros(), h.macros()):
                worst = max(worst, rel_max(a.cpu().numpy(), b))
    return worst


@pytest.mark.parametrize("coll", [0, 2])
def test_tgv_64_fp64_parity_1000_steps(coll):
    """sim_4's TGV (U0 = 0.05, tau = 0.52) on 64^3 for 1000 steps, rho and u
    every 100 steps: max |GPU - reference path| / max |reference| <= 1e-10."""
    worst = run_tgv(64, 1000, 100, coll)
    print(f"TGV 64^3 coll={coll}: worst max-rel error {worst:.3e}")
    assert worst <= 1e-10


@pytest.mark.parametrize("pattern,coll,storage,tol", [("AA", 0, "f64", 1e-13), ("AB", 0, "f64", 1e-13),
                                                      ("AA", 7, "f64", 1e-13), ("AA", 3, "f64", 1e-13),
                                                      ("AA", 0, "f32", 2e-6), ("AB", 0, "f32", 2e-6)])
def test_site_force_field_channel_matches_oracle(pattern, coll, storage, tol):
    """MACRO::computeForcing's per-site force (force_field) on a channel with
    walls, inflow and outflow, added to a homogeneous force, in the map
    kernel: populations and macros against the oracle every step."""
    X, Y, Z = 12, 9, 8
    geo = channel_geo(X, Y, Z, GEO.INFLOW, GEO.OUTFLOW_EQ if pattern == "AA" else GEO.OUTFLOW_RIGHT)
    g, h = pair(X, Y, Z, coll, storage=storage, pattern=pattern, geo=geo, nu=0.02, force=(1e-5, 0.0, -2e-5))
    rng = np.random.default_rng(17)
    F = 1e-4 * rng.standard_normal((3, X, Y, Z))
    for lat in (g, h):
        lat.set_equilibrium(1.0, 0.02, 0.0, 0.0)
        lat.set_inflow(0.02, 0.001, 0.0)
        lat.set_force_field(*F)
    for t in range(6):
        g.step(output_macro=True)
        h.step(output_macro=True)
        assert cmp_pops(g, h) < tol, (t, cmp_pops(g, h))


@pytest.mark.parametrize("storage,tol", [("f64", 1e-13), ("f32", 2e-6), ("f16dev", 3e-3)])
def test_site_force_field_without_map_runs_the_map_kernel(storage, tol):
    """A box without a map but with a force field: stepped as all PERIODIC by
    the map kernel (the map-free kernels take the homogeneous force only)."""
    X, Y, Z = 6, 70, 5
    g, h = pair(X, Y, Z, 0, storage=storage)
    r = fields(X, Y, Z, 8, amp=0.02)
    g.set_equilibrium_field(*r)
    h.set_equilibrium_field(*r)
    F = 1e-4 * np.random.default_rng(9).standard_normal((3, X, Y, Z))
    g.set_force_field(*F)
    h.set_force_field(*F)
    Wb = W[:, None, None, None] if storage == "f16dev" else 0.0
    for t in range(5):
        g.step(output_macro=True)
        h.step(output_macro=True)
        assert rel_max(g.export_f64().cpu().numpy() - Wb, h.export_f64() - Wb) < tol, t
    # the field really acts: without it the populations differ
    g2, _ = pair(X, Y, Z, 0, storage=storage)
    g2.set_equilibrium_field(*r)
    for t in range(5):
        g2.step()
    assert rel_max(g2.export_f64().cpu().numpy() - Wb, h.export_f64() - Wb) > 10 * tol


def _raw_populations(storage, n, seed):
    """Random stored values of a storage form whose decode/encode round trip is exact."""
    g = torch.Generator().manual_seed(seed)
    v = torch.rand(n, generator=g, dtype=torch.float64) * 0.2 - 0.1
    if storage == "i16dev":
        return torch.randint(-32767, 32768, (n,), generator=g, dtype=torch.int16)
    return v.to({"f64": torch.float64, "f32": torch.float32, "f32dev": torch.float32, "f16dev": torch.float16,
                 "bf16dev": torch.bfloat16}[storage])


@pytest.mark.parametrize("storage", list(STORE))
@pytest.mark.parametrize("shape", [(3, 512, 4), (4, 70, 3)])
def test_stream_only_is_free_streaming(storage, shape):
    """tnl_lbm_d3q27_stream_only runs the periodic step kernels with the
    identity collision: the even step moves slot q of each site to slot
    opp(q), the odd step then leaves population q displaced by 2 c_q, bit for
    bit (512-site rows take the 512-lane and paired shapes, 70-site rows the
    256-lane one)."""
    X, Y, Z = shape
    g = D3Q27Lattice(X, Y, Z, storage=storage, nu=0.01, store_range=0.5)
    f0 = _raw_populations(storage, g.df[0].numel(), 5)
    g.df[0].copy_(f0.to(g.df[0].device))
    s0 = g.stored().clone()
    g.stream_only()
    torch.cuda.synchronize()
    s1 = g.stored()
    bits = lambda t: t.view(torch.int16) if t.element_size() == 2 else t  # noqa: E731
    for q in range(27):
        assert torch.equal(bits(s1[OPP[q]]), bits(s0[q])), (storage, shape, "even", q)
    g.stream_only()
    torch.cuda.synchronize()
    s2 = g.stored().cpu()
    s0 = s0.cpu()
    for q in range(27):
        want = torch.roll(s0[q], shifts=(2 * int(CX[q]), 2 * int(CZ[q]), 2 * int(CY[q])), dims=(0, 1, 2))
        assert torch.equal(bits(s2[q]), bits(want)), (storage, shape, "odd", q)


@pytest.mark.parametrize("storage", ["f64", "i16dev"])
def test_turbulence_box_at_the_headline_relaxation_matches_oracle(storage):
    """The headline workload's initial condition and relaxation (seeded
    solenoidal turbulence, k0 = 8, u_rms = 0.04, tau = 0.5015, A-A periodic)
    at 32^3: FP64 per step to 1e-13 of the oracle.  The I16 form (per-class
    bounds from the field's velocity bound) is compared one step at a time
    from the same state: after each step the oracle takes the GPU's
    populations (exact multiples of the quanta, so its import re-quantises
    them to the same integers), and the next step of both must agree to one
    quantum of each slot's class, with few slots rounded apart.  Left to run
    on its own, the oracle's chain drifts away by such flips: at tau =
    0.5015 the collision scales the non-equilibrium parts by 1 - omega_1 =
    -0.994, so a one-quantum flip is carried from step to step almost
    undamped and spreads to the neighbours (8.7% of the slots after 6
    steps)."""
    from pytnl_lbm import i16_class_ranges
    from pytnl_lbm.fields import turbulence_fields
    N = 32
    r = turbulence_fields(N, 1, 8.0, 0.04)  # rho = 1, vx, vy, vz
    umax = float(np.sqrt(sum(c ** 2 for c in r[1:])).max())
    Rc = i16_class_ranges(umax) if storage == "i16dev" else None
    g, h = pair(N, N, N, storage=storage, nu=(0.5015 - 0.5) / 3.0, store_range_class=Rc)
    g.set_equilibrium_field(*r)
    h.set_equilibrium_field(*r)
    Wb = W[:, None, None, None]
    if Rc is not None:
        ncls = (CX != 0).astype(int) + (CY != 0) + (CZ != 0)
        quantum = (W * np.array(Rc)[ncls] / 32767)[:, None, None, None]
    for t in range(6):
        g.step(output_macro=True)
        h.step(output_macro=True)
        torch.cuda.synchronize()
        if storage == "f64":
            assert cmp_pops(g, h) < 1e-13, t
        else:
            gf = g.export_f64().cpu().numpy()
            diff = np.abs((gf - Wb) - (h.export_f64() - Wb))
            assert (diff <= 1.001 * quantum).all(), (t, (diff / quantum).max())
            flipped = np.mean(diff > 0.5 * quantum)
            print(f"step {t}: {100 * flipped:.3f}% of the slots one quantum apart")
            assert flipped < 0.005, t
            h.import_f64(gf)  # the next step of both starts from the GPU's state
    assert np.isfinite(g.macros()[1].cpu().numpy()).all()

