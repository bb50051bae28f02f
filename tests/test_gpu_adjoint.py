"""GPU parity of the adjoint step (tnl_lbm_d3q27_adjoint_step: the product's
hand-written reverse pass through the default cumulant) against the oracle's
forward-mode dual-number Jacobian of the same A-B step, on the same seeded
state and adjoint input: periodic with a body force, and channels with walls
at the array edges, an inflow profile and both outflows (including the
inflow-velocity gradient)."""
import numpy as np
import pytest
import torch

import oracle as O
from lbm_common import rel_max
from pytnl_lbm import D3Q27Lattice, _abi
from pytnl_lbm._abi import GEO

pytestmark = pytest.mark.gpu

TORCH = {"f64": torch.float64, "f32": torch.float32}
STORE = {"f64": _abi.STORE_F64, "f32": _abi.STORE_F32}


def channel(X, Y, Z, outflow):
    geo = np.full((X, Y, Z), GEO.FLUID, dtype=np.int16)
    geo[:, 0, :] = GEO.WALL
    geo[:, -1, :] = GEO.WALL
    geo[:, :, 0] = GEO.WALL
    geo[:, :, -1] = GEO.WALL
    geo[0, 1:-1, 1:-1] = GEO.INFLOW
    geo[-1, 1:-1, 1:-1] = outflow
    return geo


def to_dev(a, dtype):
    """(27, X, Y, Z) -> flat population layout (q, x, z, y) on the GPU."""
    return torch.as_tensor(np.ascontiguousarray(a.transpose(0, 1, 3, 2)), dtype=dtype).reshape(-1).cuda()


def from_dev(t, X, Y, Z):
    return t.double().cpu().numpy().reshape(27, X, Z, Y).transpose(0, 1, 3, 2)


def setup(case, storage, X=6, Y=70, Z=5):
    geo, force = (None, (1e-4, -2e-4, 5e-5)) if case == "periodic" else (
        channel(X, Y, Z, GEO.OUTFLOW_RIGHT if case == "outflow_right" else GEO.OUTFLOW_EQ), (0.0, 0.0, 0.0))
    mreal = _abi.REAL_F64 if storage == "f64" else _abi.REAL_F32
    g = D3Q27Lattice(X, Y, Z, storage=storage, pattern="AB", nu=0.02, geo=geo, force=force, macro_real=mreal)
    h = O.HostLattice(X, Y, Z, storage=STORE[storage], pattern=_abi.PATTERN_AB, nu=0.02, geo=geo, force=force,
                      macro_real=mreal, threads=8)
    rng = np.random.default_rng(7)
    prof = [0.03 + 0.01 * rng.standard_normal((Z, Y)), 0.005 * rng.standard_normal((Z, Y)), 0.005 * rng.standard_normal((Z, Y))]
    for lat in (g, h):
        lat.set_inflow(0.0, 0.0, 0.0, profile=prof)
    r = [1 + 0.01 * rng.standard_normal((X, Y, Z))] + list(0.03 * rng.standard_normal((3, X, Y, Z)))
    h.set_equilibrium_field(*r)
    f = h.export_f64(0) * (1 + 0.02 * rng.standard_normal((27, X, Y, Z)))
    h.import_f64(f, which=0)
    g.import_f64(h.export_f64(0), which=0)  # the stored (possibly rounded) state on both sides
    w = rng.standard_normal((27, X, Y, Z))
    return g, h, w


@pytest.mark.parametrize("case", ["periodic", "outflow_right", "outflow_eq"])
@pytest.mark.parametrize("storage,tol", [("f64", 1e-12), ("f32", 2e-5)])
def test_adjoint_step_matches_oracle(case, storage, tol):
    g, h, w = setup(case, storage)
    X, Y, Z = g.X, g.Y, g.Z
    wq = from_dev(to_dev(w, TORCH[storage]), X, Y, Z)  # the adjoint input as stored
    gh = np.zeros((3, Z, Y))
    ref = h.adjoint_step(wq, grad=gh, which=0)
    lam = torch.empty(27 * g.d.XYZ, dtype=TORCH[storage], device="cuda")
    gg = torch.zeros(3 * Y * Z, dtype=torch.float64, device="cuda")
    g.adjoint_step(to_dev(wq, TORCH[storage]), lam, gg, which=0)
    torch.cuda.synchronize()
    got = from_dev(lam, X, Y, Z)
    assert np.isfinite(got).all()
    assert rel_max(got, ref) < tol, (case, storage, rel_max(got, ref))
    if case != "periodic":
        assert rel_max(gg.cpu().numpy().reshape(3, Z, Y), gh) < tol


def test_adjoint_rejects_unsupported_configurations():
    g = D3Q27Lattice(4, 4, 4, pattern="AA", nu=0.02)
    lam = torch.zeros(27 * g.d.XYZ, dtype=torch.float64, device="cuda")
    with pytest.raises(_abi.LBMError):
        g.adjoint_step(lam, torch.zeros_like(lam))


# ---- the checkpointed, differentiable run against central differences ----------

def duct_problem(X=32, Y=16, Z=16, T=120, xp=24, storage="f64", K=None, mode="recompute", tape_f64=None):
    from pytnl_lbm import CheckpointedRun, channel_geometry, duct_flow_fields, inflow_control
    lat = D3Q27Lattice(X, Y, Z, storage=storage, pattern="AB", nu=(0.51 - 0.5) / 3, geo=channel_geometry(X, Y, Z))
    lat.set_equilibrium_field(*duct_flow_fields(X, Y, Z, 0.04))
    f0 = lat.df[0].clone()
    run = CheckpointedRun(lat, T, K, mode=mode, tape_f64=tape_f64)
    p = torch.tensor(inflow_control(Y, Z, 0.04, 0.005, seed=1), device="cuda")
    pt = torch.tensor(inflow_control(Y, Z, 0.04, 0.005, seed=2), device="cuda")
    from pytnl_lbm import density_velocity, lattice_run, populations
    with torch.no_grad():
        _, ut = density_velocity(populations(run.forward(pt, f0), lat)[:, xp])
    ut = ut.clone()

    def J(pp, ff):
        _, u = density_velocity(populations(lattice_run(run, pp, ff), lat)[:, xp])
        return 0.5 * ((u - ut) ** 2).sum()

    return lat, run, p, f0, J


@pytest.mark.parametrize("mode,tape_f64", [("recompute", None), ("segment", None), ("tape", False), ("hybrid", False),
                                           ("mixed", None)])
def test_inflow_gradient_matches_central_differences(mode, tape_f64):
    """The duct with a perturbed inlet profile, objective 1/2 sum |u - u_target|^2
    on one x plane after T steps: dJ/d(profile) . d against
    (J(p + eps d) - J(p - eps d)) / (2 eps), eps = 1e-4, on 6 random directions,
    within 1e-5 relative; and dJ/d(initial state) on 2 directions.  For each
    schedule: states recomputed, moments recorded per segment (float64),
    moments recorded for the whole run (float32), and the mixed schedule (FP64
    forward and checkpoints, FP32 recomputation and adjoint), whose FP32
    backward pass is held to the same bound against the FP64 run's
    differences."""
    lat, run, p, f0, J = duct_problem(mode=mode, tape_f64=tape_f64)
    pr, fr = p.clone().requires_grad_(), f0.clone().requires_grad_()
    Jv = J(pr, fr)
    Jv.backward()
    assert Jv.item() > 0
    gp, gf = pr.grad.clone(), fr.grad.clone()
    eps = 1e-4
    gen = torch.Generator(device="cuda").manual_seed(5)
    with torch.no_grad():
        for i in range(6):
            d = torch.randn(p.shape, generator=gen, device="cuda", dtype=torch.float64)
            d[:, 0, :] = d[:, -1, :] = 0  # wall sites carry no inflow
            d[:, :, 0] = d[:, :, -1] = 0
            d /= d.abs().max()
            fd = (J(p + eps * d, f0) - J(p - eps * d, f0)).item() / (2 * eps)
            ad = (gp * d).sum().item()
            assert abs(ad - fd) <= 1e-5 * abs(fd), (i, ad, fd)
        for i in range(2):
            d = torch.randn(f0.shape, generator=gen, device="cuda", dtype=torch.float64) * 1e-2
            fd = (J(p, f0 + eps * d) - J(p, f0 - eps * d)).item() / (2 * eps)
            ad = (gf * d).sum().item()
            assert abs(ad - fd) <= 1e-5 * abs(fd), (i, ad, fd)


def test_gradient_does_not_depend_on_the_schedule():
    """Recomputed states at three strides and float64 recordings per segment
    give the same gradient to rounding; float32 recordings of the whole run
    within 1e-6."""
    grads = {}
    for key in [("recompute", 1, None), ("recompute", None, None), ("recompute", 120, None), ("segment", None, None),
                ("segment", 7, None), ("tape", None, False), ("tape", None, True), ("hybrid", None, True)]:
        _, _, p, f0, J = duct_problem(T=120, K=key[1], mode=key[0], tape_f64=key[2])
        pr = p.clone().requires_grad_()
        J(pr, f0).backward()
        grads[key] = pr.grad.cpu().numpy()
    ref = grads[("recompute", None, None)]
    for key, g in grads.items():
        tol = 1e-6 if key == ("tape", None, False) else 1e-13
        assert rel_max(g, ref) < tol, (key, rel_max(g, ref))


def test_mixed_schedule_close_to_the_fp64_schedules():
    """The FP32 recomputation starts from an FP64 checkpoint every K steps, so
    its gradient departs from the FP64 one by FP32 rounding over at most K
    steps (and the FP32 adjoint's own rounding): ~1e-6 relative."""
    out = {}
    for mode in ("segment", "mixed"):
        _, _, p, f0, J = duct_problem(T=120, mode=mode)
        pr = p.clone().requires_grad_()
        J(pr, f0).backward()
        out[mode] = pr.grad.cpu().numpy()
    err = np.linalg.norm(out["mixed"] - out["segment"]) / np.linalg.norm(out["segment"])
    assert err < 1e-5, err


def test_backward_refuses_checkpoints_of_a_later_forward():
    """One CheckpointedRun holds the checkpoints of its latest forward pass: a
    graph whose forward was followed by another forward cannot be
    differentiated with them."""
    _, run, p, f0, J = duct_problem(T=16)
    pr = p.clone().requires_grad_()
    J1 = J(pr, f0)
    with torch.no_grad():
        J(p, f0)
    with pytest.raises(RuntimeError, match="forward again"):
        J1.backward()


def test_flat_inflow_gets_a_flat_gradient():
    from pytnl_lbm import lattice_run
    lat, run, p, f0, _ = duct_problem(T=8)
    pr = p.reshape(-1).clone().requires_grad_()
    lattice_run(run, pr, f0).sum().backward()
    assert pr.grad.shape == pr.shape


def test_ab_adjoint_differentiates_the_aa_run():
    """A-A steps from S f0 (S: pure streaming) give the A-B trajectory from
    f0, step by step; a uniform initial state is its own stream (S f0 = f0),
    so from it the A-A and A-B runs coincide and the A-B adjoint's gradient is
    the gradient of the A-A run: dJ/d(profile) of the A-B run against central
    differences of an A-A run of the same duct, J from the A-A step's macro
    output at step T (the pre-collision rho, u of that step, which the
    collision conserves).  INFLOW in, OUTFLOW_EQ out on x = X-2 with a wall
    column behind it: the zero-gradient OUTFLOW_RIGHT reads the column behind
    it after streaming and exists for the pull scheme only, and an outlet on
    the array's last plane would pull from outside the box, which A-A wraps
    (to the inlet) and A-B clamps (to itself), so the two runs would differ
    there.  The wall sends back only what it pulled from the fluid side."""
    from pytnl_lbm import CheckpointedRun, channel_geometry, density_velocity, inflow_control, lattice_run, populations
    X, Y, Z, T, xp = 32, 16, 16, 60, 24
    geo = channel_geometry(X, Y, Z, outflow=GEO.OUTFLOW_EQ)
    geo[-1] = GEO.WALL
    geo[-2, 1:-1, 1:-1] = GEO.OUTFLOW_EQ
    nu = (0.51 - 0.5) / 3
    ab = D3Q27Lattice(X, Y, Z, pattern="AB", nu=nu, geo=geo)
    ab.set_equilibrium(1.0, 0.02, 0.0, 0.0)
    f0 = ab.df[0].clone()
    run = CheckpointedRun(ab, T)
    p = torch.tensor(inflow_control(Y, Z, 0.04, 0.005, seed=1), device="cuda")
    ut = torch.full((3, Z - 2, Y - 2), 0.01, dtype=torch.float64, device="cuda")
    pr = p.clone().requires_grad_()
    # fluid sites of the plane only: the wall sites' stored populations are
    # bounced back, so their moments are not the macros of the step
    _, u = density_velocity(populations(lattice_run(run, pr, f0), ab)[:, xp, 1:-1, 1:-1])
    (0.5 * ((u - ut) ** 2).sum()).backward()

    def u_plane(pattern, pp):
        lat = D3Q27Lattice(X, Y, Z, pattern=pattern, nu=nu, geo=geo)
        lat.set_equilibrium(1.0, 0.02, 0.0, 0.0)
        lat.set_inflow(0.0, 0.0, 0.0, profile=[pp[0], pp[1], pp[2]])
        for _ in range(T - 1):
            lat.step()
        lat.step(output_macro=True)  # T is even: the A-A populations are in place again
        rho, ux, uy, uz = lat.macros()
        return torch.stack([ux[xp], uy[xp], uz[xp]]).permute(0, 2, 1)[:, 1:-1, 1:-1]

    def J_aa(pp):
        return 0.5 * ((u_plane("AA", pp) - ut) ** 2).sum().item()

    # the premise: the A-A run's macros at step T are the A-B run's
    assert (u_plane("AA", p) - u_plane("AB", p)).abs().max().item() <= 1e-14
    assert (u_plane("AB", p) - u).abs().max().item() <= 1e-13
    eps = 1e-4
    gen = torch.Generator(device="cuda").manual_seed(8)
    for _ in range(3):
        d = torch.randn(p.shape, generator=gen, device="cuda", dtype=torch.float64)
        d[:, 0, :] = d[:, -1, :] = 0
        d[:, :, 0] = d[:, :, -1] = 0
        fd = (J_aa(p + eps * d) - J_aa(p - eps * d)) / (2 * eps)
        ad = (pr.grad * d).sum().item()
        assert abs(ad - fd) <= 1e-5 * abs(fd), (ad, fd)


def test_f32_gradient_close_to_f64():
    """FP32 storage and arithmetic for the forward, the recomputation and the
    adjoint; FP64 objective and gradient accumulation."""
    out = []
    for st in ("f64", "f32"):
        _, _, p, f0, J = duct_problem(storage=st)
        pr = p.clone().requires_grad_()
        J(pr, f0).backward()
        out.append(pr.grad.cpu().numpy())
    err = np.linalg.norm(out[1] - out[0]) / np.linalg.norm(out[0])
    assert err < 1e-3, err


def test_outflow_right_inside_the_box_is_poisoned():
    """The fix-up pass covers the faces of the box; an OUTFLOW_RIGHT site
    elsewhere has no exact adjoint here, so its slots must come out NaN
    rather than silently wrong, and every other slot stays finite."""
    X, Y, Z = 8, 6, 5
    geo = np.full((X, Y, Z), GEO.PERIODIC, dtype=np.int16)
    geo[4, 2, 2] = GEO.OUTFLOW_RIGHT
    g = D3Q27Lattice(X, Y, Z, pattern="AB", nu=0.02, geo=geo)
    g.set_equilibrium(1.0, 0.01, 0.0, 0.0)
    lam_next = torch.ones(27 * g.d.XYZ, dtype=torch.float64, device="cuda")
    lam = torch.zeros_like(lam_next)
    g.adjoint_step(lam_next, lam, which=0)
    torch.cuda.synchronize()
    bad = torch.isnan(lam).sum().item()
    assert 0 < bad <= 27


# ---- the moment recording ------------------------------------------------------

@pytest.mark.parametrize("case", ["periodic", "outflow_right", "outflow_eq"])
@pytest.mark.parametrize("storage,tape_f64", [("f64", 1), ("f64", 0), ("f32", 0)])
def test_recording_step_equals_the_step_and_adjoint_from_tape(case, storage, tape_f64):
    """tnl_lbm_d3q27_step_record writes the same state as tnl_lbm_d3q27_step
    (bit for bit); its moments are those of the pulled populations; the
    adjoint from the recording equals the adjoint from the state (double
    recording: to rounding; float recording: within 1e-6)."""
    import ctypes as C
    from pytnl_lbm.lattice import _ptr, _stream_ptr
    g, _, w = setup(case, storage)
    X, Y, Z = g.X, g.Y, g.Z
    el = TORCH[storage]
    n = g.d.XYZ
    f_in = g.df[0].clone()
    g.d.df[0], g.d.df[1] = f_in.data_ptr(), g.df[1].data_ptr()
    g.d.even_iter = 1
    g._check(g.lib.tnl_lbm_d3q27_step(C.byref(g.cfg), C.byref(g.d), _abi.I64x3(0, 0, 0), _abi.I64x3(X, Y, Z), 0,
                                      _stream_ptr()))
    plain = g.df[1].clone()
    from pytnl_lbm.adjoint import tape_elements, tape_moments
    tape = torch.zeros(tape_elements(n), dtype=torch.float64 if tape_f64 else torch.float32, device="cuda")
    out = torch.zeros_like(plain)
    g.d.df[1] = out.data_ptr()
    g._check(g.lib.tnl_lbm_d3q27_step_record(C.byref(g.cfg), C.byref(g.d), _ptr(tape), tape_f64, _stream_ptr()))
    torch.cuda.synchronize()
    assert torch.equal(out, plain)
    # moments of the pulled populations at sites away from the faces (periodic neighbours there)
    if case == "periodic":
        f = from_dev(f_in, X, Y, Z)
        c = np.array([_abi_cx(), _abi_cy(), _abi_cz()])
        pulled = np.stack([np.roll(f[q], shift=(c[0, q], c[1, q], c[2, q]), axis=(0, 1, 2)) for q in range(27)])
        mom = [pulled.sum(0), *(np.tensordot(c[k], pulled, 1) for k in range(3)),
               *(np.tensordot(c[k] * c[k], pulled, 1) for k in range(3)),
               np.tensordot(c[0] * c[1], pulled, 1), np.tensordot(c[0] * c[2], pulled, 1),
               np.tensordot(c[1] * c[2], pulled, 1)]
        got = tape_moments(tape, n).double().cpu().numpy().reshape(10, X, Z, Y).transpose(0, 1, 3, 2)
        tol = 1e-14 if (storage == "f64" and tape_f64) else 2e-6  # float: 27 rounded additions
        for k in range(10):
            assert np.abs(got[k] - mom[k]).max() <= tol * max(1.0, np.abs(mom[k]).max()), k
    g._bind()
    lam_next = to_dev(w, el)
    ref = torch.empty(27 * n, dtype=el, device="cuda")
    gr_ref = torch.zeros(3 * Y * Z, dtype=torch.float64, device="cuda")
    g.d.df[0] = f_in.data_ptr()
    g._check(g.lib.tnl_lbm_d3q27_adjoint_step(C.byref(g.cfg), C.byref(g.d), _ptr(lam_next), _ptr(ref), _ptr(gr_ref),
                                              _stream_ptr()))
    got = torch.empty_like(ref)
    gr = torch.zeros_like(gr_ref)
    g._check(g.lib.tnl_lbm_d3q27_adjoint_step_recorded(C.byref(g.cfg), C.byref(g.d), _ptr(tape), tape_f64,
                                                       _ptr(lam_next), _ptr(got), _ptr(gr), _stream_ptr()))
    torch.cuda.synchronize()
    g._bind()
    # float storage: the fix-up pass adds into shared slots with atomics whose order varies, so
    # two float runs agree to float rounding only; double recordings and states agree to 1e-13
    tol = 1e-13 if (storage == "f64" and tape_f64) else 1e-6
    a, b = got.double().cpu().numpy(), ref.double().cpu().numpy()
    assert rel_max(a, b) < tol, rel_max(a, b)
    if case != "periodic":
        assert rel_max(gr.cpu().numpy(), gr_ref.cpu().numpy()) < tol


def _abi_cx():
    from pytnl_lbm.adjoint import D3Q27_CX
    return D3Q27_CX


def _abi_cy():
    from pytnl_lbm.adjoint import D3Q27_CY
    return D3Q27_CY


def _abi_cz():
    from pytnl_lbm.adjoint import D3Q27_CZ
    return D3Q27_CZ
