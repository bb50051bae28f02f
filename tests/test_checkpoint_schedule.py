"""The checkpointed run's bookkeeping (pytnl_lbm.adjoint.CheckpointedRun) on
the CPU: the lattice step is replaced by an affine map f <- A f + B p whose
adjoint is lam <- A^T lam, grad += B^T lam, so the gradient of any loss of
the final state is fixed by torch autograd through the unrolled loop.  Every
checkpoint stride K (1, sqrt T, T, non-dividing) must give that gradient, and
the forward must return exactly the unrolled final state."""
import types

import numpy as np
import pytest
import torch

from pytnl_lbm import _abi
from pytnl_lbm.adjoint import CheckpointedRun, lattice_run

Z, Y, N = 3, 4, 40  # profile 3 x Z x Y = 36 values, state of 40 values


class AffineRun(CheckpointedRun):
    def __init__(self, T, K, A, B):
        cfg = types.SimpleNamespace(pattern=_abi.PATTERN_AB, collision=_abi.COLL_CUM, storage=_abi.STORE_F64)
        d = types.SimpleNamespace(inflow_profile=[None, None, None])
        lat = types.SimpleNamespace(cfg=cfg, d=d, df=[torch.zeros(N, dtype=torch.float64)], device=torch.device("cpu"),
                                    X=1, Y=Y, Z=Z, _profiles=[None] * 3, _bind=lambda: None,
                                    set_inflow=lambda *a, **k: None)
        self.A, self.B, self.nsteps, self.nadj = A, B, 0, 0
        super().__init__(lat, T, K)

    def _step(self, src, dst, stream=None):
        assert src.data_ptr() != dst.data_ptr()
        dst.copy_(self.A @ src + self.B @ self.profile.reshape(-1))
        self.nsteps += 1

    def _adjoint(self, state, lam_next, lam, grad, stream=None):
        assert lam_next.data_ptr() != lam.data_ptr()
        lam.copy_(self.A.T @ lam_next)
        grad += (self.B.T @ lam_next).reshape(grad.shape)
        self.nadj += 1


@pytest.mark.parametrize("T,K", [(1, None), (7, 1), (7, 7), (16, None), (17, None), (23, 5)])
def test_checkpointed_gradient_equals_unrolled_autograd(T, K):
    g = torch.Generator().manual_seed(T * 31 + (K or 0))
    A = torch.eye(N, dtype=torch.float64) + 0.1 * torch.randn(N, N, generator=g, dtype=torch.float64)
    B = 0.3 * torch.randn(N, 3 * Z * Y, generator=g, dtype=torch.float64)
    p = torch.randn(3, Z, Y, generator=g, dtype=torch.float64)
    f0 = torch.randn(N, generator=g, dtype=torch.float64)
    w = torch.randn(N, generator=g, dtype=torch.float64)

    def loss(fT):
        return 0.5 * ((fT - w) ** 2).sum() + fT.sum() ** 2

    pr, fr = p.clone().requires_grad_(), f0.clone().requires_grad_()
    f = fr
    for _ in range(T):
        f = A @ f + B @ pr.reshape(-1)
    loss(f).backward()

    run = AffineRun(T, K, A, B)
    pc, fc = p.clone().requires_grad_(), f0.clone().requires_grad_()
    fT = lattice_run(run, pc, fc)
    assert torch.equal(fT.detach(), f.detach())
    nfwd = run.nsteps
    assert nfwd == T
    loss(fT).backward()
    assert run.nadj == T
    # recomputation: every step except the first of each segment
    assert run.nsteps - nfwd == T - run.nseg
    assert len(run.ckpt) == run.nseg and len(run.seg) == run.K - 1
    np.testing.assert_allclose(pc.grad.numpy(), pr.grad.numpy(), rtol=1e-12, atol=1e-12 * pr.grad.abs().max().item())
    np.testing.assert_allclose(fc.grad.numpy(), fr.grad.numpy(), rtol=1e-12, atol=1e-12 * fr.grad.abs().max().item())


def test_rejects_a_to_a_pattern():
    cfg = types.SimpleNamespace(pattern=_abi.PATTERN_AA, collision=_abi.COLL_CUM, storage=_abi.STORE_F64)
    lat = types.SimpleNamespace(cfg=cfg)
    with pytest.raises(ValueError):
        CheckpointedRun(lat, 10)
