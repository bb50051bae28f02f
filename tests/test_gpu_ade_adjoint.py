"""GPU checks of the D3Q7 adjoint (tnl_lbm_d3q7_adjoint_step) and of the
checkpointed coupled run behind torch.autograd (pytnl_lbm.CoupledScalarRun).

For a given flow state the D3Q7 part of the coupled step is affine in the
scalar populations and the source, so the oracle's forward coupled step
(oracle/lbm_oracle.cpp, kernels.h:114-196) fixes the adjoint through two
dot-product identities on the same seeded inputs:
    <lam, C delta>                  = <C^T lam, delta>
    <lam, g_out(S) - g_out(0)>      = <dJ/dS, S>
with C delta the oracle's step of the populations delta without source."""
import ctypes as C

import numpy as np
import pytest
import torch

import oracle as O
from pytnl_lbm import D3Q7Lattice, D3Q27Lattice, _abi
from pytnl_lbm._abi import ADE_GEO, GEO
from pytnl_lbm.lattice import _ptr, _stream_ptr

pytestmark = pytest.mark.gpu

ADE = {"SRT": _abi.ADE_COLL_SRT, "MRT": _abi.ADE_COLL_MRT, "CLBM": _abi.ADE_COLL_CLBM}


def geometry(kind, X, Y, Z):
    if kind == "periodic":
        return None, None
    geo = np.full((X, Y, Z), GEO.PERIODIC, dtype=np.int16)
    ageo = np.full((X, Y, Z), ADE_GEO.PERIODIC, dtype=np.int16)
    if kind == "zwalls":
        geo[:, :, 0] = geo[:, :, -1] = GEO.WALL
        ageo[:, :, 0] = ageo[:, :, -1] = ADE_GEO.WALL
    else:  # duct with inflow and outflow
        geo[:] = GEO.FLUID
        ageo[:] = ADE_GEO.FLUID
        geo[0], geo[-1] = GEO.INFLOW, GEO.OUTFLOW_RIGHT
        ageo[0], ageo[-1] = ADE_GEO.INFLOW, ADE_GEO.OUTFLOW_RIGHT
        for a in (geo,):
            a[:, 0, :] = a[:, -1, :] = a[:, :, 0] = a[:, :, -1] = GEO.WALL
        ageo[:, 0, :] = ageo[:, -1, :] = ageo[:, :, 0] = ageo[:, :, -1] = ADE_GEO.WALL
    return geo, ageo


def lay7(a):
    """(7, X, Y, Z) -> flat (q, x, z, y)."""
    return np.ascontiguousarray(np.asarray(a).transpose(0, 1, 3, 2)).reshape(-1)


def unlay7(v, X, Y, Z):
    return np.asarray(v).reshape(7, X, Z, Y).transpose(0, 1, 3, 2)


@pytest.mark.parametrize("coll", ["SRT", "MRT", "CLBM"])
@pytest.mark.parametrize("kind", ["periodic", "zwalls", "duct"])
def test_d3q7_adjoint_dot_products_against_the_oracle(kind, coll):
    X, Y, Z = 6, 70, 7
    geo, ageo = geometry(kind, X, Y, Z)
    rng = np.random.default_rng(3)
    nu, D = 0.02, 0.0125
    g = D3Q27Lattice(X, Y, Z, storage="f64", pattern="AB", nu=nu, geo=geo)
    a = D3Q7Lattice(g, collision=coll, diffusivity=D, geo=ageo)
    h = O.HostLattice(X, Y, Z, storage=_abi.STORE_F64, pattern=_abi.PATTERN_AB, nu=nu, geo=geo, threads=8)
    ha = O.HostAde(h, collision=ADE[coll], diffusivity=D, geo=ageo)
    Dfield = None
    if kind == "zwalls":  # per-site diffusivity as well
        Dfield = 0.01 + 0.01 * rng.random((X, Y, Z))
        a.set_diffusivity_field(Dfield)
        ha.set_diffusivity_field(Dfield)
    for lat in (g, h):
        lat.set_inflow(0.03, 0.01, -0.01)
    h.set_equilibrium_field(1 + 0.01 * rng.standard_normal((X, Y, Z)), *(0.04 * rng.standard_normal((3, X, Y, Z))))
    f0 = h.export_f64(0) * (1 + 0.02 * rng.standard_normal((27, X, Y, Z)))
    h.import_f64(f0, which=0)
    g.import_f64(f0, which=0)
    # the step's velocity, as the plain flow step writes it
    g.step(output_macro=True)
    torch.cuda.synchronize()
    n = g.d.XYZ
    u = g.macro[n:]

    def oracle_step(delta, S):
        h.import_f64(f0, which=0)
        h.iterations = 0
        ha.df[0][:] = lay7(delta)
        ha.set_source(S)
        h.step_coupled(ha)
        return ha.export_f64(1)

    delta = 0.1 + 0.05 * rng.standard_normal((7, X, Y, Z))
    zero = np.zeros((X, Y, Z))
    S = 1e-3 * rng.standard_normal((X, Y, Z))
    Cd = oracle_step(delta, zero)
    dS = oracle_step(delta, S) - Cd
    lam_next = rng.standard_normal((7, X, Y, Z))
    lam = torch.zeros(7 * n, dtype=torch.float64, device="cuda")
    grad = torch.zeros(n, dtype=torch.float64, device="cuda")
    a.set_source(zero)  # the adjoint's source term does not depend on the source values
    g._check(g.lib.tnl_lbm_d3q7_adjoint_step(C.byref(g.cfg), C.byref(g.d), C.byref(a.d), _ptr(u),
                                             _ptr(torch.tensor(lay7(lam_next), device="cuda")), _ptr(lam), _ptr(grad),
                                             _stream_ptr()))
    torch.cuda.synchronize()
    adj = unlay7(lam.cpu().numpy(), X, Y, Z)
    assert np.isfinite(adj).all()
    lhs, rhs = (lam_next * Cd).sum(), (adj * delta).sum()
    assert abs(lhs - rhs) <= 1e-12 * max(abs(lhs), np.abs(lam_next * Cd).sum() * 1e-3), (lhs, rhs)
    gS = grad.cpu().numpy().reshape(X, Z, Y).transpose(0, 2, 1)
    lhs, rhs = (lam_next * dS).sum(), (gS * S).sum()
    assert abs(lhs - rhs) <= 1e-9 * np.abs(lam_next * dS).sum(), (lhs, rhs)


def test_flow_step_reproduces_the_coupled_flow():
    """The backward pass recomputes the flow with the plain D3Q27 step: its
    state and velocity must equal the coupled kernel's bit for bit."""
    X, Y, Z = 8, 96, 6
    geo, ageo = geometry("zwalls", X, Y, Z)
    rng = np.random.default_rng(5)
    g = D3Q27Lattice(X, Y, Z, storage="f16dev", pattern="AB", nu=0.01, geo=geo, macro_real=_abi.REAL_F32)
    a = D3Q7Lattice(g, collision="CLBM", diffusivity=0.0125, geo=ageo, phi_ref=1.0, store_scale=100.0)
    g.set_equilibrium_field(np.ones((X, Y, Z)), *(0.03 * rng.standard_normal((3, X, Y, Z))))
    a.set_equilibrium(1.0)
    f0 = g.df[0].clone()
    g.step_coupled(a, output_macro=True)
    coupled_f, coupled_m = g.df[1].clone(), g.macro.clone()
    g.df[0].copy_(f0)
    g.iterations = 0
    g.step(output_macro=True)
    torch.cuda.synchronize()
    assert torch.equal(g.df[1], coupled_f)
    n = g.d.XYZ
    assert torch.equal(g.macro[n:], coupled_m[n:])


def scalar_problem(storage, N=(12, 10, 24), T=40, layers=6, K=None, seed=2, host=False):
    from pytnl_lbm import CoupledScalarRun, scalar_run, top_layer_variance
    X, Y, Z = N
    geo, ageo = geometry("zwalls", X, Y, Z)
    rng = np.random.default_rng(seed)
    macro = _abi.REAL_F64 if storage == "f64" else _abi.REAL_F32
    g = D3Q27Lattice(X, Y, Z, storage=storage, pattern="AB", nu=0.005 / 3, geo=geo, macro_real=macro)
    a = D3Q7Lattice(g, collision="SRT", diffusivity=0.0125, geo=ageo, phi_ref=0.0, store_scale=100.0)
    u = 0.03 * rng.standard_normal((3, X, Y, Z))
    u[:, :, :, 0] = u[:, :, :, -1] = 0
    g.set_equilibrium_field(np.ones((X, Y, Z)), *u)
    a.set_equilibrium(0.0)
    f0, g0 = g.df[0].clone(), a.df[0].clone()
    run = CoupledScalarRun(g, a, T, K, host_checkpoints=host)
    S = torch.tensor(1e-3 * rng.random(g.d.XYZ), device="cuda")

    def J(s):
        return top_layer_variance(scalar_run(run, s, g0, f0), g, layers)

    return g, run, S, J


def test_source_gradient_matches_central_differences():
    """J (the scalar variance in the top layers) is quadratic in the source,
    so central differences are exact up to rounding: FP64 adjoint vs
    (J(S + eps d) - J(S - eps d)) / (2 eps) on 4 directions, 1e-8."""
    g, run, S, J = scalar_problem("f64")
    Sr = S.clone().requires_grad_()
    Jv = J(Sr)
    Jv.backward()
    assert Jv.item() > 0
    gr = Sr.grad.clone()
    gen = torch.Generator(device="cuda").manual_seed(9)
    with torch.no_grad():
        for i in range(4):
            d = torch.randn(S.shape, generator=gen, device="cuda", dtype=torch.float64) * 1e-3
            eps = 1e-2
            fd = (J(S + eps * d) - J(S - eps * d)).item() / (2 * eps)
            ad = (gr * d).sum().item()
            assert abs(ad - fd) <= 1e-8 * abs(fd), (i, ad, fd)


def test_source_gradient_independent_of_the_stride():
    grads = []
    for K in (1, None, 40):
        _, _, S, J = scalar_problem("f64", K=K)
        Sr = S.clone().requires_grad_()
        J(Sr).backward()
        grads.append(Sr.grad.cpu().numpy())
    for gr in grads[1:]:
        assert np.abs(gr - grads[0]).max() <= 1e-13 * np.abs(grads[0]).max()


@pytest.mark.parametrize("storage", ["f64", "f16dev"])
def test_host_checkpoints_give_the_same_gradient(storage):
    """Flow checkpoints in host memory: the same states, so the same gradient
    bit for bit (K = 7 puts 6 checkpoints and a ragged last segment in play)."""
    grads = []
    for host in (False, True):
        _, run, S, J = scalar_problem(storage, K=7, host=host)
        assert run.ckpt[0].device.type == ("cpu" if host else "cuda")
        Sr = S.clone().requires_grad_()
        J(Sr).backward()
        grads.append(Sr.grad.cpu().numpy())
    assert np.array_equal(grads[0], grads[1])


@pytest.mark.parametrize("storage,tol", [("f32", 1e-5), ("f16dev", 5e-2)])
def test_reduced_precision_source_gradient_close_to_fp64(storage, tol):
    out = {}
    for st in ("f64", storage):
        _, _, S, J = scalar_problem(st)
        Sr = S.clone().requires_grad_()
        J(Sr).backward()
        out[st] = Sr.grad.cpu().numpy()
    err = np.linalg.norm(out[storage] - out["f64"]) / np.linalg.norm(out["f64"])
    assert err < tol, err
