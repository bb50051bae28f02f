"""The bookkeeping of pytnl_lbm.adjoint.CheckpointedRun, on the CPU, for its
three schedules.  The lattice step is replaced by the map
    f <- A f + B p + c * pad(f[:w]^2)
whose Jacobian at f is A + 2 c diag(pad(f[:w])): it depends on the state only
through the first w = 10 n / 27 entries, which is what a recording holds
(the ten raw moments per site of the real step).  The fake adjoint takes
that part from the state ("recompute") or from the recording ("segment",
"tape"), so a schedule that hands the adjoint the wrong step's state or
recording gives a wrong gradient.  The gradient of a loss of the final state
must equal torch autograd through the unrolled loop, for every stride K."""
import types

import numpy as np
import pytest
import torch

from pytnl_lbm import _abi
from pytnl_lbm.adjoint import CheckpointedRun, lattice_run

Z, Y, N = 3, 4, 27 * 64  # profile 3 x Z x Y = 36 values, state of 27 x 64 values
W = 640                  # the part of the state the fake records: one step's recording for 64 sites
CQ = 0.05


def fake_lattice():
    cfg = types.SimpleNamespace(pattern=_abi.PATTERN_AB, collision=_abi.COLL_CUM, storage=_abi.STORE_F64,
                                macro_real=_abi.REAL_F64)
    d = types.SimpleNamespace(inflow_profile=[None, None, None], XYZ=N // 27)
    return types.SimpleNamespace(cfg=cfg, d=d, df=[torch.zeros(N, dtype=torch.float64)], device=torch.device("cpu"),
                                 X=1, Y=Y, Z=Z, _profiles=[None] * 3, _bind=lambda: None,
                                 set_inflow=lambda *a, **k: None)


def quad(f):
    out = torch.zeros_like(f)
    out[..., :W] = CQ * f[..., :W] ** 2
    return out


class FakeRun(CheckpointedRun):
    def __init__(self, T, K, A, B, mode, budget=None):
        self.A, self.B, self.nsteps, self.nadj = A, B, 0, 0
        super().__init__(fake_lattice(), T, K, mode=mode, tape_f64=mode != "mixed", record_budget=budget)

    def _refresh_fp32(self):
        pass

    def _step(self, src, dst, tape=None, stream=None, fp32=False):
        assert src.data_ptr() != dst.data_ptr()
        assert fp32 == (src.dtype == torch.float32) == (dst.dtype == torch.float32)
        A, Bp = self.A.to(src.dtype), (self.B @ self.profile.reshape(-1)).to(src.dtype)
        if tape is not None:
            tape.copy_(src[:W])
        dst.copy_(A @ src + Bp + quad(src))
        self.nsteps += 1

    def _adjoint(self, state, lam_next, lam, grad, tape=None, stream=None, fp32=False):
        assert lam_next.data_ptr() != lam.data_ptr()
        assert fp32 == (lam.dtype == torch.float32)
        part = state[:W] if tape is None else tape
        assert (state is None) == (tape is not None)
        out = self.A.to(lam.dtype).T @ lam_next
        out[:W] += 2 * CQ * part.to(lam.dtype) * lam_next[:W]
        lam.copy_(out)
        grad += (self.B.T @ lam_next.double()).reshape(grad.shape)
        self.nadj += 1


STEP_BYTES = W * 8  # one step's float64 recording of the fake lattice


@pytest.mark.parametrize("mode,budget", [("recompute", None), ("segment", None), ("tape", None),
                                         ("hybrid", 0), ("hybrid", 12 * STEP_BYTES), ("hybrid", 1000 * STEP_BYTES),
                                         ("mixed", None)])
@pytest.mark.parametrize("T,K", [(1, None), (7, 1), (7, 7), (16, None), (17, None), (23, 5)])
def test_gradient_equals_unrolled_autograd(mode, budget, T, K):
    if mode == "mixed" and T > 20:
        pytest.skip("this map's gradient reaches 1e43 after 23 steps, past float32's range")
    g = torch.Generator().manual_seed(T * 31 + (K or 0))
    # a contraction, so that 23 steps of the quadratic map stay bounded
    A = 0.7 * torch.eye(N, dtype=torch.float64) + 0.2 / N**0.5 * torch.randn(N, N, generator=g, dtype=torch.float64)
    B = 0.05 * torch.randn(N, 3 * Z * Y, generator=g, dtype=torch.float64)
    p = torch.randn(3, Z, Y, generator=g, dtype=torch.float64)
    f0 = 0.3 * torch.randn(N, generator=g, dtype=torch.float64)
    w = torch.randn(N, generator=g, dtype=torch.float64)

    def loss(fT):
        return 0.5 * ((fT - w) ** 2).sum() + fT.sum() ** 2

    pr, fr = p.clone().requires_grad_(), f0.clone().requires_grad_()
    f = fr
    for _ in range(T):
        f = A @ f + B @ pr.reshape(-1) + quad(f)
    loss(f).backward()

    run = FakeRun(T, K, A, B, mode, budget)
    pc, fc = p.clone().requires_grad_(), f0.clone().requires_grad_()
    fT = lattice_run(run, pc, fc)
    assert torch.allclose(fT.detach(), f.detach(), rtol=1e-14, atol=1e-14)
    assert run.nsteps == T
    loss(fT).backward()
    assert run.nadj == T
    recomputed = run.nsteps - T
    if mode == "tape":
        assert recomputed == 0 and run.tape.shape[0] == T and not run.ckpt
    elif mode == "recompute":
        assert recomputed == T - run.nseg and len(run.seg) == run.K - 1
    elif mode in ("segment", "mixed"):  # every step of every segment is recorded once more
        assert recomputed == T and run.tape.shape[0] == run.K
    else:  # the recorded trailing segments are not recomputed, the others are, with recording
        assert recomputed == T - run.recorded_steps
        assert run.recorded_steps == sum(run._seg_len(j) for j in range(run.rec_first, run.nseg))
        if budget >= 1000 * STEP_BYTES:
            assert run.recorded_steps == T
        if budget == 0:
            assert run.recorded_steps == 0
    # the mixed schedule recomputes and differentiates in float32: float rounding
    tol = 1e-4 if mode == "mixed" else 1e-11
    np.testing.assert_allclose(pc.grad.numpy(), pr.grad.numpy(), rtol=tol, atol=tol * pr.grad.abs().max().item())
    np.testing.assert_allclose(fc.grad.numpy(), fr.grad.numpy(), rtol=tol, atol=tol * fr.grad.abs().max().item())
    if mode == "mixed":  # but not further off than that: float rounding of a few steps, not a wrong state
        err = (pc.grad - pr.grad).norm() / pr.grad.norm()
        assert err < 1e-5, err


def test_rejects_a_to_a_pattern_and_unknown_modes():
    cfg = types.SimpleNamespace(pattern=_abi.PATTERN_AA, collision=_abi.COLL_CUM, storage=_abi.STORE_F64)
    with pytest.raises(ValueError):
        CheckpointedRun(types.SimpleNamespace(cfg=cfg), 10)
    with pytest.raises(ValueError):
        CheckpointedRun(fake_lattice(), 10, mode="revolve")
