"""CPU baseline timing of the oracle's FP64 host path -- TEST INFRASTRUCTURE.

Run by bench.py's cpu_baseline leg as a child process, so that the OpenMP
environment (OMP_NUM_THREADS, OMP_PROC_BIND=spread, OMP_PLACES=cores) is in
place before the oracle's OpenMP runtime starts.  The oracle's host step
restates the reference's Host/OpenMP instantiation: LBMKernel per site
(kernels.h:59-112) in the loop of state.hpp:1388-1397 (collapse(2) over x and
z, y innermost), FP64, the same SoA layout.  It is a restatement, not the
reference binary (which needs TNL and network-fetched dependencies).

Samples (BASELINE.md's plan): the 64^3 Taylor-Green vortex for 1000 steps
(tau = 0.52) and a 256^3 periodic turbulence box for 20 steps (tau = 0.5015,
the 768^3 headline's workload at a size the host holds comfortably), each
with the A-B pull pattern (the reference's default) and the A-A pattern.
Each sample is the best of 3 timed runs after one untimed warm-up step; only
the step loop is timed (state.hpp:1388-1397's timer_compute).  Prints one
JSON object.
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import time
from pathlib import Path

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE))
sys.path.insert(0, str(HERE.parent / "tnl-lbm_amd"))


def host_description():
    """nproc, the CPU model, sockets and physical cores from lscpu."""
    info = {"nproc": os.cpu_count()}
    try:
        info["affinity_cpus"] = len(os.sched_getaffinity(0))
    except AttributeError:
        pass
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
        kv = {}
        for line in out.splitlines():
            if ":" in line:
                k, v = line.split(":", 1)
                kv[k.strip()] = v.strip()
        info["model"] = kv.get("Model name")
        sockets = int(kv.get("Socket(s)", "0") or 0)
        cps = int(kv.get("Core(s) per socket", "0") or 0)
        info["sockets"] = sockets
        info["physical_cores"] = sockets * cps if sockets and cps else None
        info["threads_per_core"] = int(kv.get("Thread(s) per core", "0") or 0) or None
    except (OSError, ValueError, subprocess.SubprocessError):
        pass
    return info


def run_sample(name, N, steps, pattern, nu, ic, threads, repeats):
    import numpy as np  # noqa: F401
    import oracle as O
    from pytnl_lbm import _abi
    from pytnl_lbm.fields import tgv_fields, turbulence_fields

    pat = _abi.PATTERN_AB if pattern == "AB" else _abi.PATTERN_AA
    h = O.HostLattice(N, N, N, pattern=pat, nu=nu, threads=threads)
    if ic == "tgv":
        h.set_equilibrium_field(*tgv_fields(N, N, N, 0.05))
    else:
        h.set_equilibrium_field(*turbulence_fields(N, 1, 8.0 * N / 768, 0.04))
    best = None
    for _ in range(repeats):
        h.step()  # untimed warm-up step
        t0 = time.perf_counter()
        for _ in range(steps):
            h.step()
        dt = time.perf_counter() - t0
        best = dt if best is None else min(best, dt)
    return {"sample": name, "lattice": f"{N}^3", "pattern": pattern, "steps": steps,
            "best_seconds": round(best, 3), "mlups": round(N ** 3 * steps / best / 1e6, 2), "repeats": repeats}


def main():
    p = argparse.ArgumentParser()
    p.add_argument("--threads", type=int, required=True)
    p.add_argument("--quick", action="store_true", help="small sizes (a check that the runner works)")
    a = p.parse_args()
    host = host_description()  # before OpenMP binds this thread to one place
    if a.quick:
        plan = [("tgv", 16, 20, 0.52), ("turbulence", 24, 4, 0.5015)]
        repeats = 1
    else:
        plan = [("tgv", 64, 1000, 0.52), ("turbulence", 256, 20, 0.5015)]
        repeats = 3
    samples = []
    for ic, N, steps, tau in plan:
        for pattern in ("AB", "AA"):
            name = ("64^3 Taylor-Green vortex, 1000 steps" if ic == "tgv" else
                    f"{N}^3 periodic turbulence, {steps} steps") if not a.quick else f"{ic} {N}^3"
            samples.append(run_sample(name, N, steps, pattern, (tau - 0.5) / 3.0, ic, a.threads, repeats))
            print(f"[cpu_baseline] {samples[-1]}", file=sys.stderr, flush=True)
    print(json.dumps({"host": host, "threads": a.threads,
                      "omp": {k: os.environ.get(k) for k in ("OMP_NUM_THREADS", "OMP_PROC_BIND", "OMP_PLACES")},
                      "samples": samples}), flush=True)


if __name__ == "__main__":
    main()
