#!/usr/bin/env python3
"""Benchmark of the D3Q27 cumulant A-A lattice update on one MI355X.

Headline workload (the value): FP64, 768^3 periodic box, seeded solenoidal
turbulence (E(k) ~ k^4 exp(-2 (k/8)^2), u_rms = 0.04), tau = 0.5015, A-A
pattern, default cumulant.  Beside it:
- per-precision lines (absolute FP32 and deviation storage FP32 / FP16 /
  BF16 / I16 at 512^3) with the rho/u error of each form against an FP64 run
  after 2000 steps, for a Taylor-Green vortex and for the seeded turbulence,
  both at 512^3, tau = 0.505;
- the adjoint: a 256x96x96 duct over 2000 steps (inlet-profile gradient) and
  the coupled NSE + passive scalar at 384^3 (source gradient), fwd+bwd
  against forward throughput, with central-difference checks;
- the CPU oracle (a restatement of the reference's FP64 A-B host path) timed
  on a bounded sample on the host cores.
The benchmark lattices take the library's default placement probing of
their population buffers (`placement` in the JSON line; --placement first
keeps the first allocation), and `placement_survey` times further
allocations as they land.

Prints ONE JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "tnl-lbm_amd"))

PEAK_HBM_GBS = 8000.0
METRIC = "D3Q27 A-A MLUPS & %HBM roofline per precision; adjoint fwd+bwd MLUPS"


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=500, help="timed steps of the FP64 768^3 headline (the workload is quoted on 500)")
    p.add_argument("--warmup", type=int, default=10)
    p.add_argument("--size", type=int, default=768, help="edge of the FP64 headline box")
    p.add_argument("--prec-size", type=int, default=512, help="edge of the per-precision boxes")
    p.add_argument("--prec-steps", type=int, default=100)
    p.add_argument("--prec-warmup", type=int, default=30,
                   help="untimed steps before each per-precision timing (the 2-byte odd steps take their first "
                        "10-15 launches to settle)")
    p.add_argument("--err-size", type=int, default=512, help="edge of the boxes of the error report")
    p.add_argument("--err-steps", type=int, default=2000)
    p.add_argument("--no-precision", action="store_true")
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--no-adjoint", action="store_true")
    p.add_argument("--no-shard", action="store_true")
    p.add_argument("--shard-slab", type=int, default=96, help="x extent of the one-GPU slab of the headline box")
    p.add_argument("--adjoint-steps", type=int, default=2000)
    p.add_argument("--adjoint-fd-dirs", type=int, default=16)
    p.add_argument("--coupled-steps", type=int, default=1000)
    p.add_argument("--cpu-quick", action="store_true", help="CPU baseline at small sizes (a check of the leg)")
    p.add_argument("--storage", default="f64", choices=["f64", "f32", "f32dev", "f16dev", "bf16dev", "i16dev"],
                   help="storage form of the headline case (f64 is the headline workload; the others exist "
                        "so that each kernel can be profiled through this script)")
    p.add_argument("--placement", default="auto", choices=("auto", "first"),
                   help="population buffer placement of the benchmark lattices: the library's default (auto: "
                        "buffers of 4 GiB or more are probed and a slow placement is replaced) or the first "
                        "allocation as it lands")
    p.add_argument("--allocations", type=int, default=3,
                   help="allocations of the headline box timed in all: after the headline (the first, which "
                        "gives `value`), the step and its streaming-only ceiling on allocations - 1 further "
                        "ones, each made while the previous one is still held (the achieved fraction depends "
                        "on where the driver places the allocation)")
    p.add_argument("--stride-pads", default="",
                   help="comma-separated pads of the population-plane stride, in elements: time the "
                        "headline step and its streaming-only ceiling at each stride on one allocation "
                        "(made with the largest) and report them as stride_survey")
    p.add_argument("--calibrate-copy", action="store_true",
                   help="only run a plain device copy of 2 GiB (the known-size kernel that scales the "
                        "FETCH_SIZE / WRITE_SIZE counters) and exit")
    return p.parse_args()


def progress(msg):
    """One line per finished case on stderr (stdout carries only the JSON line)."""
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def timed_steps(lat, steps, warmup, fn=None):
    """W untimed steps, then K steps bracketed by synchronisation; per-launch
    kernel durations from events recorded on the launch stream."""
    import torch

    fn = fn or lat.step
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(steps)]
    t0 = time.perf_counter()
    for i in range(steps):
        ev[i][0].record()
        fn()
        ev[i][1].record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kern = [a.elapsed_time(b) * 1e-3 for a, b in ev]
    return wall, kern


def bench_case(N, storage, steps, warmup, seed=1, u_rms=0.04, k0=8.0, tau=0.5015, store_range=1.0, ceiling=True,
               placement="auto"):
    import torch
    from pytnl_lbm import D3Q27Lattice, _abi
    from pytnl_lbm.fields import turbulence_fields_torch

    lat = D3Q27Lattice(N, N, N, storage=storage, nu=(tau - 0.5) / 3.0, store_range=store_range, placement=placement)
    u = turbulence_fields_torch(N, seed, k0, u_rms, device="cuda")
    rho = torch.ones_like(u[0])
    lat.set_equilibrium_field(rho, *u)
    del u, rho
    torch.cuda.empty_cache()
    wall, kern = timed_steps(lat, steps, warmup)
    bytes_per_site = 2 * 27 * _abi.STORAGE_BYTES[lat.storage]
    sites = N ** 3
    avg_kernel = sum(kern) / len(kern)
    res = {
        "mlups": sites * steps / wall / 1e6,
        "ms_per_step": wall / steps * 1e3,
        "kernel_ms_avg": avg_kernel * 1e3,
        "kernel_ms_even": sum(kern[0::2]) / max(1, len(kern[0::2])) * 1e3,
        "kernel_ms_odd": sum(kern[1::2]) / max(1, len(kern[1::2])) * 1e3,
        "bytes_per_site": bytes_per_site,
        "achieved_GBs": bytes_per_site * sites / avg_kernel / 1e9,
    }
    res["roofline_frac"] = res["achieved_GBs"] / PEAK_HBM_GBS
    if lat.placement is not None:
        res["placement"] = {"policy": placement, "probed_frac": [round(r / PEAK_HBM_GBS, 4)
                                                                 for r in lat.placement["probed_GBs"]],
                            "kept": lat.placement["kept"]}
    # FP64 reductions of the final state (mass, NaN check)
    lat.compute_macro()
    red = lat.reduce_macro()
    res["mass_rel_drift"] = abs(red[0] / sites - 1.0)
    res["nan_sites"] = red[2]
    if ceiling:
        # the same kernels, launch shape and allocation with the identity
        # collision: the memory ceiling of the step's access pattern here
        _, ck = timed_steps(lat, steps, 2, fn=lat.stream_only)
        c_avg = sum(ck) / len(ck)
        res["ceiling"] = {
            "what": "tnl_lbm_d3q27_stream_only: the step kernels with the identity collision (free streaming), "
                    "same lattice allocation, same bytes",
            "kernel_ms_avg": round(c_avg * 1e3, 3),
            "kernel_ms_even": round(sum(ck[0::2]) / max(1, len(ck[0::2])) * 1e3, 3),
            "kernel_ms_odd": round(sum(ck[1::2]) / max(1, len(ck[1::2])) * 1e3, 3),
            "roofline_frac": round(bytes_per_site * sites / c_avg / 1e9 / PEAK_HBM_GBS, 4),
            "step_rate_over_ceiling_rate": round(c_avg / avg_kernel, 4)}
    del lat
    torch.cuda.empty_cache()
    return res


def shard_case(N, slab, steps, warmup, tau=0.5015, placement="auto"):
    """One GPU's share of the FP64 N^3 box split into x-slabs over N/slab
    GPUs: slab x N x N sites plus a ghost plane per side, stepped with
    DISTRIBUTED_X.  The slab is its own periodic neighbour, so the halo
    exchange (tnl_lbm_d3q27_halo_x: PULL after the even step, PUSH after the
    odd one) runs as device-local plane copies, standing in for the xGMI
    transfers of a real split.  Times the step kernel and the exchange
    separately with events on the launch stream."""
    import torch
    from pytnl_lbm import D3Q27Lattice, _abi

    lat = D3Q27Lattice(slab, N, N, storage="f64", nu=(tau - 0.5) / 3.0, overlap=(1, 0, 0), placement=placement)
    lat.set_equilibrium(1.0, 0.02, -0.01, 0.015)
    ev = [[torch.cuda.Event(enable_timing=True) for _ in range(3)] for _ in range(steps)]
    for i in range(warmup + steps):
        e = ev[i - warmup] if i >= warmup else None
        if e:
            e[0].record()
        lat.step(flags=_abi.FLAG_DISTRIBUTED_X)
        if e:
            e[1].record()
        lat.halo_x(lat, _abi.HALO_PULL if lat.iterations % 2 == 1 else _abi.HALO_PUSH)
        if e:
            e[2].record()
    torch.cuda.synchronize()
    k = sum(a.elapsed_time(b) for a, b, _ in ev) / steps * 1e-3
    h = sum(b.elapsed_time(c) for _, b, c in ev) / steps * 1e-3
    sites = slab * N * N
    face = 9 * N * N * 8  # bytes per face per exchange
    res = {"workload": f"x-slab {slab}x{N}x{N} + 1 ghost plane per side (one GPU's share of {N}^3 on {N // slab} GPUs), "
                       "FP64, A-A, DISTRIBUTED_X",
           "mlups": round(sites / k / 1e6, 1), "kernel_ms_avg": round(k * 1e3, 3),
           "roofline_frac": round(432 * sites / k / 1e9 / PEAK_HBM_GBS, 4),
           "halo_ms_avg": round(h * 1e3, 4), "halo_bytes_per_step": 2 * face,
           "halo_note": "device-local copies of the 2 x 9 boundary planes; over one xGMI link per face "
                        f"(~153 GB/s) a face would take ~{face / 153e9 * 1e3:.2f} ms, overlappable with the interior"}
    if lat.placement is not None:
        res["placement"] = {"probed_frac": [round(r / PEAK_HBM_GBS, 4) for r in lat.placement["probed_GBs"]],
                            "kept": lat.placement["kept"]}
    del lat
    torch.cuda.empty_cache()
    return res


def placement_survey(N, storage, count, steps=6, tau=0.5015):
    """The headline step and the streaming-only step on `count` successive
    allocations of the N^3 box, each kept as allocated (no placement probing)
    and made while the previous one is held (so the driver places it
    elsewhere); uniform initial state.  Per allocation: the placement probe's
    rate right after the allocation, then the fraction of 8 TB/s of both
    steps, and their ratio."""
    import torch
    from pytnl_lbm import D3Q27Lattice, _abi

    out, held = [], None
    for i in range(count):
        lat = D3Q27Lattice(N, N, N, storage=storage, nu=(tau - 0.5) / 3.0, placement="first")
        probe = lat._probe_rate(lat.df[0]) / PEAK_HBM_GBS
        lat.set_equilibrium(1.0, 0.02, -0.01, 0.015)
        _, k = timed_steps(lat, steps, 2)
        _, c = timed_steps(lat, steps, 2, fn=lat.stream_only)
        gb = 2 * 27 * _abi.STORAGE_BYTES[lat.storage] * N ** 3 / 1e9
        ks, cs = sum(k) / len(k), sum(c) / len(c)
        out.append({"allocation": i + 1, "probe_frac": round(probe, 4), "step_frac": round(gb / ks / PEAK_HBM_GBS, 4),
                    "ceiling_frac": round(gb / cs / PEAK_HBM_GBS, 4), "step_rate_over_ceiling_rate": round(cs / ks, 4)})
        progress(f"allocation {i + 1}: probe {out[-1]['probe_frac']}, step {out[-1]['step_frac']}, "
                 f"ceiling {out[-1]['ceiling_frac']}")
        del held
        torch.cuda.empty_cache()
        held = lat
        del lat
    del held
    torch.cuda.empty_cache()
    return out


def stride_survey(N, storage, pads, steps=8, rounds=2, tau=0.5015):
    """The headline step and its streaming-only ceiling at several plane
    strides XYZ = N^3 + pad on one allocation: the lattice is allocated with
    the largest stride and its data struct takes each smaller one in turn
    (every slot stays inside the allocation), so that the allocation's
    placement is the same for all.  Uniform initial state."""
    import torch
    from pytnl_lbm import D3Q27Lattice, _abi

    base = N ** 3
    lat = D3Q27Lattice(N, N, N, storage=storage, nu=(tau - 0.5) / 3.0, XYZ=base + max(pads))
    lat.set_equilibrium(1.0, 0.02, -0.01, 0.015)
    el = _abi.STORAGE_BYTES[lat.storage]
    gb = 2 * 27 * el * base / 1e9
    out = []
    for r in range(rounds):
        for pad in pads:
            lat.d.XYZ = base + pad
            _, c = timed_steps(lat, steps, 2, fn=lat.stream_only)
            _, k = timed_steps(lat, steps, 2)
            ks, cs = sum(k) / len(k), sum(c) / len(c)
            out.append({"round": r, "pad_bytes": pad * el, "step_frac": round(gb / ks / PEAK_HBM_GBS, 4),
                        "ceiling_frac": round(gb / cs / PEAK_HBM_GBS, 4)})
            progress(f"stride pad {pad * el} B: step {out[-1]['step_frac']}, ceiling {out[-1]['ceiling_frac']}")
    del lat
    torch.cuda.empty_cache()
    return out


def precision_errors(N, steps, storages, ic="tgv"):
    """The same initial condition run in FP64 and in each storage form; L2
    and max error of rho and u against the FP64 run after `steps` steps.
    ic "tgv": Taylor-Green vortex, U0 = 0.05; "turbulence": the seeded
    spectrum of the headline box (u_rms = 0.04, k0 = 8).  tau = 0.505."""
    import torch
    from pytnl_lbm import D3Q27Lattice, tgv_fields
    from pytnl_lbm.fields import turbulence_fields_torch

    from pytnl_lbm import i16_class_ranges

    nu = (0.505 - 0.5) / 3.0
    if ic == "tgv":
        u_max = 0.05
    else:
        u = turbulence_fields_torch(N, 1, 8.0, 0.04, device="cuda")
        u_max = float(torch.sqrt(u[0] ** 2 + u[1] ** 2 + u[2] ** 2).max())
        del u
    # I16: one bound per direction class from the flow's velocity bound
    Rc = i16_class_ranges(u_max)

    def init(lat):
        if ic == "tgv":
            lat.set_equilibrium_field(*tgv_fields(N, N, N, 0.05))
        else:
            u = turbulence_fields_torch(N, 1, 8.0, 0.04, device="cuda")
            lat.set_equilibrium_field(torch.ones_like(u[0]), *u)
            del u

    def run(storage):
        lat = D3Q27Lattice(N, N, N, storage=storage, nu=nu, store_range=1.0, store_range_class=Rc)
        init(lat)
        torch.cuda.empty_cache()
        for t in range(steps):
            lat.step(output_macro=(t == steps - 1))
        torch.cuda.synchronize()
        out = [m.to(torch.float64).clone() for m in lat.macros()]
        del lat
        torch.cuda.empty_cache()
        return out

    ref = run("f64")
    res = {}
    for st in storages:
        got = run(st)
        e = {}
        d = got[0] - ref[0]
        e["rho_L2"] = float(torch.linalg.vector_norm(d) / torch.linalg.vector_norm(ref[0] - 1.0))
        e["rho_max"] = float(d.abs().max() / (ref[0] - 1.0).abs().max())
        du = torch.stack([got[i] - ref[i] for i in (1, 2, 3)])
        ub = torch.stack(ref[1:])
        e["u_L2"] = float(torch.linalg.vector_norm(du) / torch.linalg.vector_norm(ub))
        e["u_max"] = float(du.abs().max() / ub.abs().max())
        res[st] = e
        del got, du, ub
    return res, {"case": ic, "lattice": f"{N}^3", "tau": 0.505, "steps": steps,
                 "i16_store_range_class": [round(r, 5) for r in Rc], "u_max": round(u_max, 5),
                 "U0" if ic == "tgv" else "u_rms": 0.05 if ic == "tgv" else 0.04,
                 "rho_norm": "relative to rho-1 of the FP64 run"}


def adjoint_case(X, Y, Z, T, xp, storage="f64", fd_dirs=16, eps=1e-4, adj_reps=20, mode="segment"):
    """Duct 256x96x96 (walls in y, z; equilibrium inlet at x = 0, zero-gradient
    outlet at x = X-1), tau = 0.51, developed-flow initial state.  Control: the
    inlet profile (parabolic, Umax = 0.04, plus a seeded smooth perturbation of
    amplitude 0.005); objective J = 1/2 sum |u - u_target|^2 in FP64 on the
    x = xp plane after T steps, the target from a second seed.  Times the
    forward A-B step alone and the forward + checkpointed backward through
    torch.autograd (the second such pass: the steady state of an optimisation
    loop), and checks dJ/dp against central differences."""
    import torch
    from pytnl_lbm import (CheckpointedRun, D3Q27Lattice, _abi, channel_geometry, density_velocity,
                           duct_flow_fields, inflow_control, lattice_run, populations)

    torch.cuda.empty_cache()
    lat = D3Q27Lattice(X, Y, Z, storage=storage, pattern="AB", nu=(0.51 - 0.5) / 3, geo=channel_geometry(X, Y, Z))
    lat.set_equilibrium_field(*duct_flow_fields(X, Y, Z, 0.04))
    f0 = lat.df[0].clone()
    run = CheckpointedRun(lat, T, mode=mode)
    p = torch.tensor(inflow_control(Y, Z, 0.04, 0.005, seed=1), device="cuda")
    pt = torch.tensor(inflow_control(Y, Z, 0.04, 0.005, seed=2), device="cuda")
    with torch.no_grad():
        _, ut = density_velocity(populations(run.forward(pt, f0), lat)[:, xp])
        ut = ut.clone()

    def J(pp):
        _, u = density_velocity(populations(lattice_run(run, pp, f0), lat)[:, xp])
        return 0.5 * ((u - ut) ** 2).sum()

    sites = X * Y * Z
    # the plain forward: T A-B steps, no checkpoints, no recording
    run.profile.copy_(p)
    src, dst = run.pp
    src.copy_(f0)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(T):
        run._step(src, dst)
        src, dst = dst, src
    torch.cuda.synchronize()
    t_fwd = time.perf_counter() - t0
    lat._bind()
    # forward + backward, after one untimed pass that brings the caching allocator to its steady state
    # (autograd's full-lattice gradient buffers are fresh device allocations on the first pass only)
    pw = p.clone().requires_grad_()
    J(pw).backward()
    del pw
    pr = p.clone().requires_grad_()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    Jv = J(pr)
    Jv.backward()
    torch.cuda.synchronize()
    t_fb = time.perf_counter() - t0
    g = pr.grad.clone()
    # adjoint kernel alone (the form this schedule uses), on one state or recording
    fp32 = mode == "mixed"  # its adjoint runs over float fields and float recordings
    lam = [torch.randn(f0.numel(), device="cuda", dtype=torch.float32 if fp32 else f0.dtype) * 1e-3 for _ in range(2)]
    grad = torch.zeros(3, Z, Y, dtype=torch.float64, device="cuda")
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(adj_reps)]
    tape = run.rec[0] if run.rec is not None else (None if run.tape is None else run.tape[0])
    for i in range(adj_reps + 2):
        if i >= 2:
            ev[i - 2][0].record()
        run._adjoint(run.ckpt[0] if tape is None else None, lam[i % 2], lam[1 - i % 2], grad, tape, fp32=fp32)
        if i >= 2:
            ev[i - 2][1].record()
    torch.cuda.synchronize()
    lat._bind()
    adj_s = sum(a.elapsed_time(b) for a, b in ev) / adj_reps * 1e-3
    el = 4 if fp32 else _abi.STORAGE_BYTES[lat.storage]
    # lambda_next read + lambda written + the Jacobian's input (27 pulled populations or 10 recorded moments)
    jac = 27 * el if tape is None else 10 * tape.element_size()
    adj_site = 2 * 27 * el + jac
    adj_bytes = adj_site * sites
    res = {
        "workload": f"duct {X}x{Y}x{Z}, tau=0.51, T={T}, J on plane x={xp}, {storage.upper()} A-B cumulant",
        "schedule": mode if tape is None else f"{mode} ({'float64' if run.tape_f64 else 'float32'} moment recording)",
        "schedule_note": {"hybrid": "sqrt(T) checkpoints; the forward records the moments of the trailing segments "
                                    "that fit in 80% of free HBM, only the rest is recomputed (a full recording "
                                    "when it fits: not a sqrt(T) memory budget)",
                          "segment": "sqrt(T) checkpoints; each segment recomputed with recording",
                          "recompute": "sqrt(T) checkpoints; each segment's states recomputed and kept",
                          "tape": "every step recorded, no checkpoints (not a sqrt(T) memory budget)",
                          "mixed": "sqrt(T) FP64 checkpoints of the FP64 forward; each segment recomputed in FP32 "
                                   "with float recording, adjoint over float fields in FP64 arithmetic, FP64 gradient accumulation"}[mode],
        "sqrt_T_memory": mode in ("segment", "recompute", "mixed"),
        "memory_GB": round(run.memory_bytes() / 1e9, 2),
        "forward_mlups": round(sites * T / t_fwd / 1e6, 1),
        "fwd_bwd_mlups": round(sites * T / t_fb / 1e6, 1),
        "ratio": round(t_fwd / t_fb, 4),
        "ratio_definition": "fwd_bwd_mlups / forward_mlups; forward = the plain A-B step, T steps; fwd+bwd = the second autograd pass (the first warms the allocator)",
        "checkpoint_stride": run.K if mode != "tape" else None,
        "checkpoints": run.nseg if mode != "tape" else 0,
        "recorded_steps": run.recorded_steps if mode == "hybrid" else (T if mode == "tape" else 0),
        "J": Jv.item(),
        "adjoint_kernel": {"ms_avg": round(adj_s * 1e3, 3), "bytes_per_site": adj_site,
                           "achieved_GBs": round(adj_bytes / adj_s / 1e9, 1),
                           "roofline_frac": round(adj_bytes / adj_s / 1e9 / PEAK_HBM_GBS, 4)},
    }
    if fd_dirs:
        gen = torch.Generator(device="cuda").manual_seed(11)
        worst = 0.0
        with torch.no_grad():
            for _ in range(fd_dirs):
                d = torch.randn(p.shape, generator=gen, device="cuda", dtype=torch.float64)
                d[:, 0, :] = d[:, -1, :] = 0
                d[:, :, 0] = d[:, :, -1] = 0
                d /= d.abs().max()
                fd = (J(p + eps * d) - J(p - eps * d)).item() / (2 * eps)
                worst = max(worst, abs((g * d).sum().item() - fd) / abs(fd))
        res["fd_check"] = {"eps": eps, "directions": fd_dirs, "max_rel_err": float(f"{worst:.3e}")}
    del run, lat, J, Jv, pr, g, lam, grad, tape
    torch.cuda.empty_cache()
    return res


def coupled_adjoint_case(N, T, storage, layers=32, fd_dirs=0, grad_out=False, K=None, host_checkpoints=False,
                         fb_warmup=1):
    """Coupled NSE + D3Q7 passive scalar on N^3, periodic in x and y, walls
    in z; seeded turbulence (u_rms = 0.03, k0 = 6; zero on the walls),
    tau = 0.505; D3Q7 CLBM with tau_ADE = 0.55 (D = 0.0125); source: 16
    Gaussian blobs, sigma = 6, amplitude 1e-3.  Objective: FP64 variance of
    the scalar over the top `layers` z-layers after T steps; gradient with
    respect to the source through the checkpointed run, timed after
    `fb_warmup` untimed passes (the first one allocates autograd's buffers)."""
    import numpy as np
    import torch
    from pytnl_lbm import (ADE_GEO, GEO, CoupledScalarRun, D3Q7Lattice, D3Q27Lattice, _abi, gaussian_blobs,
                           scalar_run, to_storage, top_layer_variance)
    from pytnl_lbm.fields import turbulence_fields_torch

    geo = np.full((N, N, N), GEO.PERIODIC, dtype=np.int16)
    geo[:, :, 0] = geo[:, :, -1] = GEO.WALL
    ageo = np.full((N, N, N), ADE_GEO.PERIODIC, dtype=np.int16)
    ageo[:, :, 0] = ageo[:, :, -1] = ADE_GEO.WALL
    macro = _abi.REAL_F64 if storage == "f64" else _abi.REAL_F32
    lat = D3Q27Lattice(N, N, N, storage=storage, pattern="AB", nu=(0.505 - 0.5) / 3, geo=geo, macro_real=macro)
    ade = D3Q7Lattice(lat, collision="CLBM", diffusivity=(0.55 - 0.5) / 4, geo=ageo, phi_ref=0.0, store_scale=1000.0)
    u = turbulence_fields_torch(N, 7, 6.0, 0.03, device="cuda")
    for c in u:
        c[:, :, 0] = 0
        c[:, :, -1] = 0
    lat.set_equilibrium_field(torch.ones_like(u[0]), *u)
    del u
    ade.set_equilibrium(0.0)
    f0, g0 = lat.df[0].clone(), ade.df[0].clone()
    src = to_storage(torch.tensor(gaussian_blobs(N, N, N, 16, 6.0, 1e-3, seed=11))).reshape(-1).to("cuda")
    torch.cuda.empty_cache()
    run = CoupledScalarRun(lat, ade, T, K=K, host_checkpoints=host_checkpoints)
    sites = N ** 3

    def J(s):
        return top_layer_variance(scalar_run(run, s, g0, f0), lat, layers)

    torch.cuda.synchronize()
    t0 = time.perf_counter()
    run.forward(src, g0, f0)
    torch.cuda.synchronize()
    t_fwd = time.perf_counter() - t0
    for _ in range(fb_warmup):
        sw = src.clone().requires_grad_()
        J(sw).backward()
        del sw
    sr = src.clone().requires_grad_()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    Jv = J(sr)
    Jv.backward()
    torch.cuda.synchronize()
    t_fb = time.perf_counter() - t0
    res = {
        "workload": f"coupled NSE + D3Q7 CLBM {N}^3, {storage.upper()}, periodic x,y, walls z, T={T}, "
                    f"J = variance of phi over the top {layers} layers",
        "forward_mlups": round(sites * T / t_fwd / 1e6, 1),
        "fwd_bwd_mlups": round(sites * T / t_fb / 1e6, 1),
        "ratio": round(t_fwd / t_fb, 4),
        "fwd_bwd_timed_pass": fb_warmup + 1,
        "checkpoint_stride": run.K,
        "checkpoints_in": "host memory" if host_checkpoints else "HBM",
        "J": Jv.item(),
    }
    grad = sr.grad.clone()
    if fd_dirs:
        gen = torch.Generator(device="cuda").manual_seed(13)
        worst = 0.0
        with torch.no_grad():
            for _ in range(fd_dirs):
                d = torch.randn(src.shape, generator=gen, device="cuda", dtype=torch.float64) * 1e-3
                eps = 1e-2
                fd = (J(src + eps * d) - J(src - eps * d)).item() / (2 * eps)
                worst = max(worst, abs((grad * d).sum().item() - fd) / abs(fd))
        res["fd_check"] = {"eps": 1e-2, "directions": fd_dirs, "max_rel_err": float(f"{worst:.3e}")}
    del run, ade, lat
    torch.cuda.empty_cache()
    return (res, grad) if grad_out else res


def cpu_baseline(quick=False):
    """The oracle's FP64 host step (a restatement of the reference's
    Host/OpenMP path: LBMKernel in the loop of state.hpp:1388-1397) on every
    CPU this job is given, in a child process that starts with
    OMP_PROC_BIND=spread and OMP_PLACES=cores: the 64^3 Taylor-Green vortex
    for 1000 steps and a 256^3 turbulence box for 20 steps, A-B and A-A, best
    of 3 (oracle/cpu_baseline.py).  The value is the 256^3 A-A sample."""
    import subprocess

    cpus = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    env_threads = os.environ.get("OMP_NUM_THREADS")
    threads = min(cpus, int(env_threads)) if env_threads and env_threads.isdigit() else cpus
    env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="spread", OMP_PLACES="cores")
    cmd = [sys.executable, str(ROOT / "oracle" / "cpu_baseline.py"), "--threads", str(threads)]
    if quick:
        cmd.append("--quick")
    r = subprocess.run(cmd, env=env, capture_output=True, text=True, check=True)
    res = json.loads(r.stdout.strip().splitlines()[-1])
    head = [s for s in res["samples"] if s["pattern"] == "AA"][-1]
    host = res["host"]
    return {"value": head["mlups"], "unit": "MLUPS", "cores": threads, "kind": "port",
            "sample": f"oracle FP64 D3Q27 cumulant (restatement of the reference Host/OpenMP path), "
                      f"{head['sample']}, A-A, best of {head['repeats']}, {threads} OpenMP threads "
                      f"(OMP_PROC_BIND=spread, OMP_PLACES=cores)",
            "host": {"model": host.get("model"), "sockets": host.get("sockets"),
                     "physical_cores": host.get("physical_cores"), "nproc": host.get("nproc"),
                     "cpus_given_to_this_job": cpus},
            "samples": res["samples"]}


def calibrate_copy(nbytes=2 << 30, reps=5):
    """A plain device-to-device copy of known size: the FETCH_SIZE and
    WRITE_SIZE counters read for it give the scale from counter units to
    bytes."""
    import torch

    a = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    b.copy_(a)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        b.copy_(a)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / reps
    return {"copy_bytes_read": nbytes, "copy_bytes_written": nbytes, "reps": reps + 1,
            "GBs": 2 * nbytes / dt / 1e9}


def measured_traffic(workload):
    """HBM bytes per launch of the dominant kernel on this workload, from the
    committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary (scaled by the plain
    copy), or None when no summary covers the workload."""
    path = ROOT / "profiles" / "hbm_traffic.json"
    if not path.exists():
        return None
    try:
        table = json.loads(path.read_text())
    except ValueError:
        return None
    row = table.get("workloads", {}).get(workload)
    if not row:
        return None
    return row.get("bytes_per_launch")


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); multi-GPU sharding is "
              "described in DESIGN.md and not measured here", file=sys.stderr)
        sys.exit(2)
    import torch

    torch.cuda.init()
    if args.calibrate_copy:
        print(json.dumps(calibrate_copy()), flush=True)
        return
    N = args.size
    st = args.storage
    tau = 0.5015 if st == "f64" else 0.505
    head = bench_case(N, st, args.steps, args.warmup, tau=tau, store_range=0.5 if st == "i16dev" else 1.0,
                      placement=args.placement)
    kernel = (f"k_aa_even_pair<{st.upper()},CUM> (even) / k_aa_periodic<{st.upper()},CUM> (odd)"
              if st in ("f16dev", "bf16dev", "i16dev") else f"k_aa_periodic<{st.upper()},CUM> (even/odd)")
    workload = f"D3Q27 cumulant A-A {st.upper()}, {N}^3 periodic box, tau={tau}"
    out = {
        "metric": METRIC,
        "value": round(head["mlups"], 1),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(head["ms_per_step"], 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if st == "f64" else "f32",
        "data": "synthetic: seeded solenoidal turbulence, E(k)~k^4 exp(-2(k/8)^2), u_rms=0.04, rho=1",
        "config": {"workload": workload, "lattice": [N, N, N], "collision": "CUM", "pattern": "AA", "storage": st},
        "roofline": {"bound": "hbm", "achieved": round(head["achieved_GBs"], 1), "peak": PEAK_HBM_GBS,
                     "unit": "GB/s", "frac": round(head["roofline_frac"], 4), "traffic": measured_traffic(workload),
                     "kernel": kernel, "bytes_per_site": head["bytes_per_site"],
                     "algorithmic_bytes_per_launch": head["bytes_per_site"] * N ** 3,
                     "kernel_ms_avg": round(head["kernel_ms_avg"], 3),
                     "kernel_ms_even": round(head["kernel_ms_even"], 3),
                     "kernel_ms_odd": round(head["kernel_ms_odd"], 3)},
        "checks": {"mass_rel_drift": head["mass_rel_drift"], "nan_sites": head["nan_sites"]},
        "ceiling": head.get("ceiling"),
        "placement": head.get("placement"),
    }
    progress(f"{workload}: {head['mlups']:.0f} MLUPS, {100 * head['roofline_frac']:.1f}% of 8 TB/s")
    if args.allocations > 1:
        out["placement_survey"] = placement_survey(N, st, args.allocations - 1)
    if args.stride_pads:
        out["stride_survey"] = stride_survey(N, st, [int(v) for v in args.stride_pads.split(",")])
    if not args.no_shard and st == "f64" and N % args.shard_slab == 0:
        out["shard"] = shard_case(N, args.shard_slab, min(args.steps, 40), 2, placement=args.placement)
        progress(f"shard {out['shard']['workload']}: {out['shard']['mlups']:.0f} MLUPS, "
                 f"halo {out['shard']['halo_ms_avg']} ms per step")
    if not args.no_precision:
        per = {}
        P = args.prec_size
        for pst in ("f32", "f32dev", "f16dev", "bf16dev", "i16dev"):
            r = bench_case(P, pst, args.prec_steps, args.prec_warmup, tau=0.505,
                           store_range=0.5 if pst == "i16dev" else 1.0, placement=args.placement)
            per[f"{pst}_{P}"] = {"mlups": round(r["mlups"], 1), "achieved_GBs": round(r["achieved_GBs"], 1),
                                 "roofline_frac": round(r["roofline_frac"], 4), "bytes_per_site": r["bytes_per_site"],
                                 "kernel_ms_avg": round(r["kernel_ms_avg"], 3),
                                 "kernel_ms_even": round(r["kernel_ms_even"], 3),
                                 "kernel_ms_odd": round(r["kernel_ms_odd"], 3),
                                 "ceiling": r.get("ceiling"), "placement": r.get("placement")}
            progress(f"{pst} {P}^3: {r['mlups']:.0f} MLUPS, {100 * r['roofline_frac']:.1f}% of 8 TB/s")
        spec = []
        for ic in ("tgv", "turbulence"):
            errs, sp = precision_errors(args.err_size, args.err_steps, ("f32", "f32dev", "f16dev", "bf16dev", "i16dev"),
                                        ic=ic)
            spec.append(sp)
            progress(f"precision errors ({ic}, {args.err_size}^3, {args.err_steps} steps) done")
            for pst, e in errs.items():
                per[f"{pst}_{P}"].setdefault("error_vs_fp64", {})[ic] = {k: float(f"{v:.3e}") for k, v in e.items()}
        out["per_precision"] = per
        out["precision_error_case"] = spec
    if not args.no_adjoint:
        adj = {}
        for st, mode in (("f64", "mixed"), ("f64", "segment"), ("f64", "recompute"), ("f32", "segment"),
                         ("f64", "hybrid")):
            adj[f"{st}_{mode}"] = adjoint_case(256, 96, 96, args.adjoint_steps, 200, storage=st, mode=mode,
                                               fd_dirs=args.adjoint_fd_dirs if (st, mode) == ("f64", "mixed") else 0)
            progress(f"adjoint {st} {mode}: ratio {adj[f'{st}_{mode}'].get('ratio')}")
        # coupled scalar at 384^3: throughput in FP16 storage; its gradient against the FP64 adjoint of the
        # same run (FP64 flow checkpoints in host memory: 10 x 12.2 GB)
        r16, g16 = coupled_adjoint_case(384, args.coupled_steps, "f16dev", grad_out=True)
        adj["coupled_f16dev_384"] = r16
        progress("coupled adjoint 384^3 F16 done")
        g16 = g16.cpu()
        r64, g64 = coupled_adjoint_case(384, args.coupled_steps, "f64", grad_out=True, K=100, host_checkpoints=True,
                                        fb_warmup=0)
        g64 = g64.cpu()
        adj["coupled_f64_384_reference"] = r64
        adj["coupled_f16dev_384"]["grad_error_vs_fp64"] = {
            "lattice": "384^3",
            "rel_L2": float(f"{float(torch.linalg.vector_norm(g16 - g64) / torch.linalg.vector_norm(g64)):.3e}"),
            "rel_max": float(f"{float((g16 - g64).abs().max() / g64.abs().max()):.3e}")}
        del g16, g64
        progress("coupled adjoint 384^3 FP64 reference done")
        out["adjoint"] = adj
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(args.cpu_quick)
        progress(f"cpu baseline: {out['cpu_baseline']['value']} MLUPS on {out['cpu_baseline']['cores']} threads")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
