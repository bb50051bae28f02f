"""Initial-condition fields (host, FP64 numpy), uploaded by the caller.

Generated on the host in double precision so that every backend starts from
bit-identical (rho, u) arrays.
"""
from __future__ import annotations

import numpy as np


def tgv_coordinates(N: int) -> np.ndarray:
    """Cell-centred coordinates of sim_4's lattice (sim_4.cu:340-348,
    lattice.h:71-74): x_i = origin + (i - 1/2) dl with dl = 2 pi / N and
    origin = -dl N / 2 + dl, i.e. x_i = dl (i + 1/2) - pi."""
    dl = 2 * np.pi / N
    origin = -dl / 2 * N + dl
    return origin + (np.arange(N, dtype=np.float64) - 0.5) * dl


def tgv_fields(X: int, Y: int, Z: int, U0: float):
    """Taylor-Green vortex of sim_4.cu:145-163 with L = 1 and V_0 = U0 in
    lattice units: u = U0 sin x cos y cos z, v = -U0 cos x sin y cos z, w = 0,
    rho = 1 + 3 (U0^2/16)(cos 2x + cos 2y)(cos 2z + 2).  Returns [x, y, z]
    arrays (rho, vx, vy, vz)."""
    x = tgv_coordinates(X)[:, None, None]
    y = tgv_coordinates(Y)[None, :, None]
    z = tgv_coordinates(Z)[None, None, :]
    u = U0 * np.sin(x) * np.cos(y) * np.cos(z)
    v = -U0 * np.cos(x) * np.sin(y) * np.cos(z)
    w = np.zeros((X, Y, Z))
    rho = 1 + 3 * (U0 * U0 / 16) * (np.cos(2 * x) + np.cos(2 * y)) * (np.cos(2 * z) + 2)
    shape = (X, Y, Z)
    return (np.broadcast_to(rho, shape).copy(), np.broadcast_to(u, shape).copy(),
            np.broadcast_to(v, shape).copy(), w)


def turbulence_fields(N: int, seed: int, k0: float, u_rms: float, chunk_z: int | None = None):
    """Seeded solenoidal random velocity on an N^3 periodic box with energy
    spectrum E(k) ~ k^4 exp(-2 (k/k0)^2) (integer box wavenumbers), zero mean,
    Nyquist modes removed, scaled to u_rms = sqrt(<|u|^2>/3); rho = 1.
    Returns float64 [x, y, z] arrays (rho, vx, vy, vz)."""
    rng = np.random.default_rng(seed)
    k = np.fft.fftfreq(N, d=1.0 / N)
    kz = np.fft.rfftfreq(N, d=1.0 / N)
    KX, KY, KZ = np.meshgrid(k, k, kz, indexing="ij")
    K2 = KX * KX + KY * KY + KZ * KZ
    Kmag = np.sqrt(K2)
    amp = np.zeros_like(Kmag)
    nz = Kmag > 0
    E = Kmag[nz] ** 4 * np.exp(-2.0 * (Kmag[nz] / k0) ** 2)
    amp[nz] = np.sqrt(E / (4 * np.pi * K2[nz]))
    nyq = (np.abs(KX) == N // 2) | (np.abs(KY) == N // 2) | (np.abs(KZ) == N // 2)
    amp[nyq] = 0.0
    uh = []
    for _ in range(3):
        re = rng.standard_normal(Kmag.shape)
        im = rng.standard_normal(Kmag.shape)
        uh.append((re + 1j * im) * amp)
    del amp, re, im
    # solenoidal projection u <- u - k (k.u)/|k|^2
    K2s = np.where(nz, K2, 1.0)
    div = (KX * uh[0] + KY * uh[1] + KZ * uh[2]) / K2s
    uh[0] -= KX * div
    uh[1] -= KY * div
    uh[2] -= KZ * div
    del div
    u = [np.fft.irfftn(h, s=(N, N, N), axes=(0, 1, 2)) for h in uh]
    del uh
    for c in u:
        c -= c.mean()
    rms = np.sqrt((u[0] ** 2 + u[1] ** 2 + u[2] ** 2).mean() / 3.0)
    scale = u_rms / rms
    for c in u:
        c *= scale
    rho = np.ones((N, N, N))
    return rho, u[0], u[1], u[2]


def turbulence_fields_torch(N: int, seed: int, k0: float, u_rms: float, device="cuda", dtype=None):
    """turbulence_fields() on a torch device (same spectrum, projection and
    normalisation; the random stream is torch's).  Returns [x, y, z] float64
    tensors (vx, vy, vz) on `device`; rho = 1 is implied."""
    import torch

    dt = dtype or torch.float64
    g = torch.Generator(device=device)
    g.manual_seed(seed)
    k = torch.fft.fftfreq(N, d=1.0 / N, device=device, dtype=dt)
    kz = torch.fft.rfftfreq(N, d=1.0 / N, device=device, dtype=dt)
    KX, KY, KZ = k[:, None, None], k[None, :, None], kz[None, None, :]
    K2 = KX * KX + KY * KY + KZ * KZ
    Kmag = torch.sqrt(K2)
    E = Kmag ** 4 * torch.exp(-2.0 * (Kmag / k0) ** 2)
    amp = torch.sqrt(E / (4 * torch.pi * torch.where(K2 > 0, K2, torch.ones_like(K2))))
    amp[0, 0, 0] = 0
    half = N // 2
    amp[half, :, :] = 0
    amp[:, half, :] = 0
    amp[:, :, -1] = 0 if N % 2 == 0 else amp[:, :, -1]
    del E, Kmag
    shape = amp.shape
    uh = []
    for _ in range(3):
        re = torch.randn(shape, generator=g, device=device, dtype=dt)
        im = torch.randn(shape, generator=g, device=device, dtype=dt)
        uh.append(torch.complex(re * amp, im * amp))
        del re, im
    del amp
    K2s = torch.where(K2 > 0, K2, torch.ones_like(K2))
    div = (KX * uh[0] + KY * uh[1] + KZ * uh[2]) / K2s
    uh[0] -= KX * div
    uh[1] -= KY * div
    uh[2] -= KZ * div
    del div, K2s, K2
    u = []
    for h in uh:
        u.append(torch.fft.irfftn(h, s=(N, N, N), dim=(0, 1, 2)))
    del uh
    for c in u:
        c -= c.mean()
    rms = torch.sqrt((u[0] ** 2 + u[1] ** 2 + u[2] ** 2).mean() / 3.0)
    s = u_rms / rms
    return [c * s for c in u]


def channel_geometry(X: int, Y: int, Z: int, outflow: int | None = None) -> np.ndarray:
    """Square duct along x: bounce-back WALL sites on the y and z edges, the
    equilibrium velocity inlet (INFLOW) on the x = 0 plane and the
    zero-gradient outlet (OUTFLOW_RIGHT, bc.h:355-371) on x = X-1, as the
    reference's channel setups place them (sim_adjoint.cu:51-66 and
    setBoundaryX/Y/Z in lbm_block.hpp).  Returns an int16 [x, y, z] array of
    GEO codes."""
    from ._abi import GEO
    geo = np.full((X, Y, Z), GEO.FLUID, dtype=np.int16)
    geo[0, :, :] = GEO.INFLOW
    geo[-1, :, :] = GEO.OUTFLOW_RIGHT if outflow is None else outflow
    geo[:, 0, :] = GEO.WALL
    geo[:, -1, :] = GEO.WALL
    geo[:, :, 0] = GEO.WALL
    geo[:, :, -1] = GEO.WALL
    return geo


def duct_profile(Y: int, Z: int, umax: float) -> np.ndarray:
    """Parabolic x-velocity of the duct, umax at the centre and zero on the wall
    sites y = 0, Y-1 and z = 0, Z-1; shape (Z, Y)."""
    y = np.arange(Y, dtype=np.float64) / (Y - 1)
    z = np.arange(Z, dtype=np.float64) / (Z - 1)
    return umax * 16.0 * np.outer(z * (1 - z), y * (1 - y))


def smooth_perturbation(Y: int, Z: int, amplitude: float, seed: int, modes: int = 4) -> np.ndarray:
    """Seeded smooth random field of shape (3, Z, Y): for each velocity
    component a sum of sin(m pi y') sin(n pi z') modes (m, n <= modes,
    coefficients N(0,1)/(m^2+n^2)) that vanishes on the walls, scaled to
    max |.| = amplitude."""
    rng = np.random.default_rng(seed)
    y = np.arange(Y, dtype=np.float64) / (Y - 1)
    z = np.arange(Z, dtype=np.float64) / (Z - 1)
    out = np.zeros((3, Z, Y))
    for k in range(3):
        for m in range(1, modes + 1):
            for n in range(1, modes + 1):
                a = rng.standard_normal() / (m * m + n * n)
                out[k] += a * np.outer(np.sin(n * np.pi * z), np.sin(m * np.pi * y))
        out[k] *= amplitude / np.abs(out[k]).max()
    return out


def inflow_control(Y: int, Z: int, umax: float, amplitude: float, seed: int) -> np.ndarray:
    """Inlet velocity profile (3, Z, Y): the duct profile in x plus a seeded
    smooth perturbation in every component."""
    p = smooth_perturbation(Y, Z, amplitude, seed)
    p[0] += duct_profile(Y, Z, umax)
    return p


def duct_flow_fields(X: int, Y: int, Z: int, umax: float):
    """Developed duct flow as initial condition: rho = 1, u = the parabolic
    duct profile on every x plane.  Returns [x, y, z] arrays (rho, vx, vy, vz)."""
    vx = np.broadcast_to(duct_profile(Y, Z, umax).T[None], (X, Y, Z)).copy()
    zero = np.zeros((X, Y, Z))
    return np.ones((X, Y, Z)), vx, zero, zero.copy()


def gaussian_blobs(X: int, Y: int, Z: int, count: int, sigma: float, amplitude: float, seed: int,
                   periodic=(True, True, False)) -> np.ndarray:
    """Sum of `count` seeded Gaussian blobs amplitude * exp(-r^2 / (2 sigma^2))
    with centres uniform in the box (periodic distance along the periodic
    axes), each evaluated within 5 sigma of its centre.  Returns an [x, y, z]
    float64 array."""
    rng = np.random.default_rng(seed)
    out = np.zeros((X, Y, Z))
    dims = (X, Y, Z)
    r = int(np.ceil(5 * sigma))
    for _ in range(count):
        c = rng.random(3) * np.array(dims)
        idx, fac = [], []
        for a in range(3):
            n = dims[a]
            lo, hi = int(np.floor(c[a])) - r, int(np.floor(c[a])) + r + 1
            pts = np.arange(lo, hi)
            if periodic[a]:
                ix = np.mod(pts, n)
            else:
                keep = (pts >= 0) & (pts < n)
                pts, ix = pts[keep], pts[keep]
            d = pts - c[a]
            idx.append(ix)
            fac.append(np.exp(-d * d / (2 * sigma * sigma)))
        blob = amplitude * fac[0][:, None, None] * fac[1][None, :, None] * fac[2][None, None, :]
        np.add.at(out, np.ix_(idx[0], idx[1], idx[2]), blob)
    return out
