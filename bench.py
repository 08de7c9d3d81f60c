"""Benchmark: FP64 cell-updates/s of the 2D Navier-Stokes + photochemistry
CRWENO5 RK4 step on one MI355X.

Workload (the configuration the headline metric is quoted on): 1024 x 1024
periodic grid, nvars = 11 (4 flow + 7 species), CRWENO5 component-wise
reconstruction (one tridiagonal system per grid line, variable and
reconstruction), Roe upwinding, fourth-order viscous terms (Re = 1000 set
through the total pressure), O2/O3 photochemistry with a Gaussian beam map,
RK4 with dt = 0.01.  A step is one RK4 time step (four right-hand sides);
inputs are resident in HBM before timing starts.

Prints one JSON line (see README of the benchmark contract in DESIGN.md).
"""
from __future__ import annotations

import argparse
import json
import os
import re
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "FP64 cell-updates/s (2D NS CRWENO5); batched-serve p99 latency vs serial"
HBM_PEAK_GBS = 8000.0
# algorithmic bytes per cell-update of a stage-fused classical RK step:
# per stage read stage state + u_n, read/write the accumulator, write the next
# stage state = 5 * 8 * nvars bytes, times the number of stages
STEP_BYTES_PER_CELL = lambda nv, ns: 40 * nv * ns  # noqa: E731


def kernel_algorithmic_bytes(name, nx, ny, nv, g, nstages, crweno):
    """Compulsory bytes of one launch of a kernel class (its inputs read once
    and outputs written once).  One launch covers both sweeps (x and y) of an
    nx x ny grid; per sweep the counts below are for `lines` lines of N
    interior points (the two sweeps have the same counts on a square grid)."""
    lines, N = ny, nx
    pos = N + 2 * g
    sweeps = 2
    if name == "recon":   # u and f at every position of every line; uL uR fL fR
        return sweeps * 8 * (2 * nv * pos * lines + 4 * nv * (N + 1) * lines)
    if name == "upwind":  # uL uR fL fR + 4 flow vars at the two cells; fI
        return sweeps * 8 * (4 * nv * (N + 1) * lines + 2 * 4 * (N + 1) * lines + nv * (N + 1) * lines)
    if name == "flux":
        return sweeps * 8 * 2 * nv * pos * lines
    if name == "stage_bc":  # u + up to one Udot read, U and UT written (ghosted)
        return 8 * nv * (nx + 2 * g) * (ny + 2 * g) * 4
    return None


def cpu_baseline(tmp, nx_sample=256, steps=10):
    """Reference PIAFS (built from source, serial, 1 core) on a bounded sample of
    the same workload: the config deck at nx_sample^2 with the same dx, dt and
    physics.  Falls back to the CPU restatement where the reference binary is
    absent."""
    from piafs_amd import decks
    dk = decks.ns2d_photochem(nx=nx_sample, n_iter=steps, dx=10.0 * np.pi / 1024)
    d = os.path.join(tmp, "cpu")
    dk.write(d)
    ref = os.path.join(ROOT, "oracle", "_ref", "PIAFS")
    cells = nx_sample * nx_sample * steps
    sample = (f"config-2 deck restricted to {nx_sample}x{nx_sample} (same dx, dt, physics), "
              f"{steps} RK4 steps")
    if os.path.exists(ref):
        r = subprocess.run([ref], cwd=d, capture_output=True, text=True, env=dict(os.environ, OMP_NUM_THREADS="1"))
        m = re.search(r"Solver runtime \(in seconds\):\s*(\S+)", r.stdout)
        if r.returncode == 0 and m:
            secs = float(m.group(1))
            return {"value": cells / secs, "unit": "cell-updates/s", "cores": 1, "kind": "reference",
                    "sample": sample + "; reference binary built from source (oracle/_ref/PIAFS), serial",
                    "seconds": secs}
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import _oracle as orc
    c = orc.OCase(d)
    t0 = time.perf_counter()
    c.step(steps)
    secs = time.perf_counter() - t0
    return {"value": cells / secs, "unit": "cell-updates/s", "cores": 1, "kind": "port",
            "sample": sample + "; CPU restatement (oracle/liboracle.so), serial", "seconds": secs}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--nx", type=int, default=1024)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-profile", action="store_true")
    ap.add_argument("--parity", action="store_true", help="check the final state against the reference fixture")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MB",
                    help="only run a plain device-to-device copy of MB megabytes (10 times), the reference "
                         "kernel for scaling the FETCH_SIZE / WRITE_SIZE counters")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        n = args.copy_calibration * 1000 * 1000 // 8
        a = torch.ones(n, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        for _ in range(10):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": 8 * n, "copy_bytes_written": 8 * n, "launches": 10}), flush=True)
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this build runs on one GPU only (nothing here uses a second GPU); "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import piafs_amd
    from piafs_amd import decks

    nx = args.nx
    with tempfile.TemporaryDirectory() as tmp:
        dk = decks.ns2d_photochem(nx=nx, n_iter=args.warmup + args.steps, dx=10.0 * np.pi / 1024 if nx != 1024 else None)
        d = os.path.join(tmp, "case")
        dk.write(d)
        case = piafs_amd.Case(d, device=0)
        inf = case.info
        nv, g = inf.nvars, inf.ghosts
        case.step(args.warmup)
        case.sync()
        case.profile(not args.no_profile)
        t0 = time.perf_counter()
        case.step(args.steps)
        case.sync()
        secs = time.perf_counter() - t0
        ktimes = {} if args.no_profile else case.kernel_times()
        case.profile(False)
        cells = nx * nx
        value = cells * args.steps / secs
        line = {
            "metric": METRIC, "value": value, "unit": "cell-updates/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": 1e3 * secs / args.steps, "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
            "config": {"workload": f"2D NavierStokes + photochemistry {nx}x{nx}, nvars {nv}, CRWENO5 components, "
                                   f"Roe, 4th-order viscous (Re 1000), RK4 dt 0.01",
                       "grid": [nx, nx], "nvars": nv, "steps": args.steps, "parallelism": "single GPU"},
        }
        if ktimes:
            # per-class average launch time (ms) x launches per step -> share of the step
            ns = inf.nstages
            per_step = {"stage_bc": ns, "flux": ns, "recon": ns, "upwind": ns, "qderiv": ns, "viscflux": ns,
                        "assemble": ns, "update": 1}
            share = {k: v * per_step.get(k, 1) for k, v in ktimes.items()}
            dom = max(share.items(), key=lambda kv: kv[1])[0]
            byts = kernel_algorithmic_bytes(dom, nx, nx, nv, g, inf.nstages, True)
            ms = ktimes[dom]
            ach = byts / (ms * 1e-3) / 1e9 if byts else None
            line["roofline"] = {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                "frac": (ach / HBM_PEAK_GBS) if ach else None, "traffic": None,
                                "kernel": dom, "kernel_ms": ms, "algorithmic_bytes_per_launch": byts}
            line["kernel_ms"] = ktimes
            line["kernel_ms_per_step"] = share
        step_gbs = STEP_BYTES_PER_CELL(nv, inf.nstages) * value / 1e9
        line["step_roofline"] = {"bytes_per_cell_update": STEP_BYTES_PER_CELL(nv, inf.nstages),
                                 "achieved": step_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                 "frac": step_gbs / HBM_PEAK_GBS}
        if args.parity and nx == 1024 and args.warmup + args.steps == 200:
            fx = os.path.join(ROOT, "tests", "golden", "config2_1024_ref.npz")
            if os.path.exists(fx):
                z = np.load(fx)
                U = case.solution().reshape(nx, nx, nv)
                sub = U[::32, ::32, :].reshape(-1, nv)
                ref = z["sub"].reshape(-1, nv)
                err = np.sqrt(((sub - ref) ** 2).sum(0)) / np.sqrt((ref ** 2).sum(0))
                line["parity"] = {"vs": "reference run (oracle/_ref) 1024^2 x 200 steps, stride-32 subsample",
                                  "max_rel_l2_per_var": float(err.max())}
        case.close()
        if not args.no_cpu_baseline:
            line["cpu_baseline"] = cpu_baseline(tmp)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
