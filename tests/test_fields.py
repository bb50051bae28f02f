"""The synthetic initial conditions of the measured workloads (CPU): the
seeded solenoidal turbulence of the 768^3 / 512^3 / 384^3 boxes and the
Taylor-Green vortex of sim_4.cu:126-171.

turbulence_fields (numpy) and turbulence_fields_torch (the generator the
benchmark uses on the GPU, run here on the CPU device) must give: u_rms =
sqrt(<|u|^2>/3) equal to the request, zero mean, a divergence-free field
(k . u_hat = 0 for every integer wavenumber), no Nyquist content, and a
shell-summed energy spectrum proportional to k^4 exp(-2 (k/k0)^2)."""
import numpy as np
import pytest
import torch

from pytnl_lbm.fields import tgv_fields, turbulence_fields, turbulence_fields_torch


def spectral_checks(u, N, k0, u_rms):
    u = np.stack([np.asarray(c, dtype=np.float64) for c in u])
    assert abs(np.sqrt((u ** 2).sum(0).mean() / 3) / u_rms - 1) < 1e-12
    assert np.abs(u.mean(axis=(1, 2, 3))).max() < 1e-15
    uh = np.fft.fftn(u, axes=(1, 2, 3))
    k = np.fft.fftfreq(N, d=1.0 / N)
    K = np.stack(np.meshgrid(k, k, k, indexing="ij"))
    Kmag = np.sqrt((K ** 2).sum(0))
    # sum |k . u_hat|^2 against sum |k|^2 |u_hat|^2: rounding only (modes in the
    # spectrum's far tail are themselves at rounding level, so no per-mode ratio)
    div2 = (np.abs((K * uh).sum(0)) ** 2).sum()
    assert div2 / (Kmag ** 2 * (np.abs(uh) ** 2).sum(0)).sum() < 1e-26
    nyq = (np.abs(K) == N // 2).any(axis=0)
    assert np.abs(uh[:, nyq]).max() < 1e-12 * np.abs(uh).max()
    # shell spectrum E(k) = 1/2 sum_{k-1/2 < |kappa| <= k+1/2} |u_hat|^2 against the model shape
    E = 0.5 * (np.abs(uh) ** 2).sum(0)
    shells = np.arange(3, int(2.5 * k0) + 1)
    meas = np.array([E[(Kmag > s - 0.5) & (Kmag <= s + 0.5)].sum() for s in shells])
    model = shells ** 4 * np.exp(-2 * (shells / k0) ** 2)
    ratio = meas / model
    assert np.abs(ratio / np.median(ratio) - 1).max() < 0.35, ratio / np.median(ratio)
    # the shape, not just the scale: the peak shell sits at k0 / sqrt(2) * sqrt(2) = k0 within a shell or two
    peak = shells[np.argmax(meas)]
    assert abs(peak - k0) <= 2


@pytest.mark.parametrize("N,k0,u_rms,seed", [(48, 8.0, 0.04, 1), (40, 6.0, 0.03, 7)])
def test_numpy_turbulence(N, k0, u_rms, seed):
    rho, vx, vy, vz = turbulence_fields(N, seed, k0, u_rms)
    assert (rho == 1).all()
    spectral_checks((vx, vy, vz), N, k0, u_rms)
    again = turbulence_fields(N, seed, k0, u_rms)
    assert np.array_equal(again[1], vx)  # seeded


@pytest.mark.parametrize("N,k0,u_rms,seed", [(48, 8.0, 0.04, 1), (40, 6.0, 0.03, 7)])
def test_torch_turbulence(N, k0, u_rms, seed):
    u = turbulence_fields_torch(N, seed, k0, u_rms, device="cpu")
    assert all(c.dtype == torch.float64 and c.shape == (N, N, N) for c in u)
    spectral_checks([c.numpy() for c in u], N, k0, u_rms)


def test_taylor_green_initial_condition():
    """sim_4.cu:145-163 with the cell-centred coordinates of sim_4.cu:340:
    u = U0 (sin x cos y cos z, -cos x sin y cos z, 0),
    rho = 1 + 3 (U0^2/16)(cos 2x + cos 2y)(cos 2z + 2); divergence-free."""
    N, U0 = 16, 0.05
    rho, vx, vy, vz = tgv_fields(N, N, N, U0)
    x = 2 * np.pi * (np.arange(N) + 0.5) / N - np.pi
    X, Y, Zc = np.meshgrid(x, x, x, indexing="ij")
    assert np.abs(vx - U0 * np.sin(X) * np.cos(Y) * np.cos(Zc)).max() < 1e-15
    assert np.abs(vy + U0 * np.cos(X) * np.sin(Y) * np.cos(Zc)).max() < 1e-15
    assert np.abs(vz).max() == 0
    assert np.abs(rho - 1 - 3 * U0 ** 2 / 16 * (np.cos(2 * X) + np.cos(2 * Y)) * (np.cos(2 * Zc) + 2)).max() < 1e-15
