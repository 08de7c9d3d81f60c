"""Benchmark: FP64 cell-updates/s of the 2D Navier-Stokes + photochemistry
CRWENO5 RK4 step on one MI355X.

Workload (the configuration the headline metric is quoted on): 1024 x 1024
periodic grid, nvars = 11 (4 flow + 7 species), CRWENO5 component-wise
reconstruction (one tridiagonal system per grid line, variable and
reconstruction), Roe upwinding, fourth-order viscous terms (Re = 1000 set
through the total pressure), O2/O3 photochemistry with a Gaussian beam map,
RK4 with dt = 0.01.  A step is one RK4 time step (four right-hand sides);
inputs are resident in HBM before timing starts.

Prints one JSON line: metric, value, unit, n_gpus, steps, warmup, ms_per_step, the
workload, the roofline of the dominant kernel and the CPU baseline.
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


def species_fused(nx, ny, nv, paired):
    """Whether the library's reconstruction writes the species' Roe fluxes
    itself: NS2D with species, paired Roe, and 32- or 64-lane line teams in
    both sweeps (lines of 258 .. 1025 rows)."""
    return paired and nv > 4 and 258 <= nx <= 1025 and 258 <= ny <= 1025


def kernel_algorithmic_bytes(name, nx, ny, nv, g, nstages, crweno, paired=True):
    """Compulsory bytes of one launch of a kernel class (its inputs read once
    and outputs written once).  One launch covers both sweeps (x and y) of an
    nx x ny grid; per sweep the counts below are for `lines` lines of N
    interior points (the two sweeps have the same counts on a square grid).
    paired: component-wise interpolation with Roe or Rusanov, where the
    reconstruction writes only the two combinations 0.5 (uR - uL) and
    0.5 (fL + fR) that the upwinding reads, instead of uL, uR, fL and fR."""
    lines, N = ny, nx
    pos = N + 2 * g
    sweeps = 2
    nio = 2 if paired else 4
    fused = species_fused(nx, ny, nv, paired)
    if name in ("recon", "recon_full"):  # u and f at every position of every line; the interface arrays
        if fused:  # flow: the two combinations; species: their flux, from |v_n| (one value per interface)
            return sweeps * 8 * (2 * nv * pos * lines + (2 * 4 + (nv - 4) + 1) * (N + 1) * lines)
        return sweeps * 8 * (2 * nv * pos * lines + nio * nv * (N + 1) * lines)
    if name == "upwind":  # the interface arrays + 4 flow vars at the two cells; fI
        if fused:  # the flow variables only
            return sweeps * 8 * (nio * 4 * (N + 1) * lines + 2 * 4 * (N + 1) * lines + 4 * (N + 1) * lines)
        return sweeps * 8 * (nio * nv * (N + 1) * lines + 2 * 4 * (N + 1) * lines + nv * (N + 1) * lines)
    if name == "flux":  # the state and the flux at every position; |v_n| per interface when fused
        return sweeps * 8 * (2 * nv * pos * lines + ((N + 1) * lines if fused else 0))
    if name == "stage_bc":  # u + up to one Udot read, U and UT written (ghosted)
        return 8 * nv * (nx + 2 * g) * (ny + 2 * g) * 4
    return None


def measured_traffic(nx, kernel):
    """HBM bytes per launch of a kernel class from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE runs of this workload (scaled by a copy kernel of
    known size), or None when no profile of this grid and kernel is committed."""
    path = os.path.join(ROOT, "profiles", f"ns2d_crweno5_{nx}_step_kernels.md")
    if not os.path.exists(path):
        return None
    m = re.search(r"`k_%s`: HBM traffic ([0-9.]+) MB read \+ ([0-9.]+) MB written" % re.escape(kernel),
                  open(path).read())
    return (float(m.group(1)) + float(m.group(2))) * 1e6 if m else None


# MI355X: 256 CUs x 4 SIMDs at 2.4 GHz; a wave64 FP64 VALU instruction issues
# every 4 cycles on a 16-lane SIMD (the rate behind the 78.6 TFLOP/s FP64
# vector peak: 256 * 4 * 16 lanes * 2 flops * 2.4e9)
WAVE_INSTS_PER_S = 256 * 4 * 2.4e9 / 4
FP64_PEAK_FLOPS = 78.6e12


def valu_counters(nx, kernel):
    """Per-launch VALU counters of `kernel` from the committed rocprofv3 --pmc
    runs of this workload (the `counters-json` line of the profile summary),
    or None."""
    path = os.path.join(ROOT, "profiles", f"ns2d_crweno5_{nx}_step_kernels.md")
    if not os.path.exists(path):
        return None
    m = re.search(r"^counters-json: (\{.*\})\s*$", open(path).read(), re.M)
    if not m:
        return None
    return json.loads(m.group(1)).get("k_" + kernel)


def valu_fractions(c, ms):
    """VALU issue fraction (every VALU wave-instruction costs one FP64 issue
    slot), FP64 flop fraction of the vector peak (64 lanes per wave-instruction,
    FMA = 2 flops) and the fraction of the attainable roofline
    min(HBM BW x AI, FP64 peak), AI from the same flops over the kernel's
    measured HBM bytes."""
    t = ms * 1e-3
    out = {"valu_wave_insts_per_launch": c["SQ_INSTS_VALU"],
           "valu_issue_frac": c["SQ_INSTS_VALU"] / WAVE_INSTS_PER_S / t}
    f64 = ("SQ_INSTS_VALU_ADD_F64", "SQ_INSTS_VALU_MUL_F64", "SQ_INSTS_VALU_FMA_F64", "SQ_INSTS_VALU_TRANS_F64")
    if all(k in c for k in f64):
        flops = 64.0 * (c[f64[0]] + c[f64[1]] + 2 * c[f64[2]] + c[f64[3]])
        out["fp64_flops_per_launch"] = flops
        out["fp64_frac_of_peak"] = flops / t / FP64_PEAK_FLOPS
        out["fp64_insts_frac_of_valu"] = sum(c[k] for k in f64) / c["SQ_INSTS_VALU"]
        hbm = c.get("hbm_bytes")
        if hbm:
            attainable = min(HBM_PEAK_GBS * 1e9 * flops / hbm, FP64_PEAK_FLOPS)
            out["attainable_flops"] = attainable
            out["attainable_frac"] = flops / t / attainable
    return out


def cpu_model():
    """The host CPU's model name (/proc/cpuinfo) and its logical core count."""
    name = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                name = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return name, os.cpu_count()


def cpu_baseline(tmp, nx_sample=256, steps=10):
    """Reference PIAFS (built from source, serial, 1 core) on a bounded sample of
    the same workload: the benchmark deck at nx_sample^2 with the same dx, dt and
    physics.  Falls back to the CPU restatement where the reference binary is
    absent."""
    from piafs_amd import decks
    dk = decks.ns2d_photochem(nx=nx_sample, n_iter=steps, dx=10.0 * np.pi / 1024)
    d = os.path.join(tmp, "cpu")
    dk.write(d)
    ref = os.path.join(ROOT, "oracle", "_ref", "PIAFS")
    cells = nx_sample * nx_sample * steps
    model, ncpu = cpu_model()
    sample = (f"the 1024^2 chemistry deck restricted to {nx_sample}x{nx_sample} (same dx, dt, physics), "
              f"{steps} RK4 steps, on one core of the host CPU ({model}, {ncpu} logical cores)")
    if os.path.exists(ref):
        r = subprocess.run([ref], cwd=d, capture_output=True, text=True, env=dict(os.environ, OMP_NUM_THREADS="1"))
        m = re.search(r"Solver runtime \(in seconds\):\s*(\S+)", r.stdout)
        if r.returncode == 0 and m:
            secs = float(m.group(1))
            return {"value": cells / secs, "unit": "cell-updates/s", "cores": 1, "kind": "reference",
                    "sample": sample + "; reference binary built from source (oracle/_ref/PIAFS), serial",
                    "seconds": secs, "cpu_model": model}
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import _oracle as orc
    c = orc.OCase(d)
    t0 = time.perf_counter()
    c.step(steps)
    secs = time.perf_counter() - t0
    return {"value": cells / secs, "unit": "cell-updates/s", "cores": 1, "kind": "port",
            "sample": sample + "; CPU restatement (oracle/liboracle.so), serial", "seconds": secs,
            "cpu_model": model}


def _write_members(tmp, members, sub):
    dirs = []
    for m in members:
        d = os.path.join(tmp, sub, m.name)
        m.deck.write(d)
        dirs.append(d)
    return dirs


def bench_batch(tmp, n_cases=256, steps=20, serial_steps=10, oracle_check=True):
    """Ensemble of independent 2D NS + photochemistry cases (sides log-uniform
    in [128, 512], Re log-uniform in [1e2, 1e4], beam amplitude in [0.5, 2],
    WENO5/CRWENO5 50/50): the whole ensemble advanced by the batched launches
    (one per kernel class and RK stage, replayed from a HIP graph) against the
    same cases solved one at a time through the single-case path."""
    import piafs_amd
    from piafs_amd import decks
    warm = 1
    members = decks.ensemble(n_cases, n_iter=warm + steps)
    t0 = time.perf_counter()
    dirs = _write_members(tmp, members, "ens")
    t_write = time.perf_counter() - t0
    cells = sum(m.cells for m in members)
    b = piafs_amd.Batch(dirs)
    tuned_rows, _ = b.autotune()
    b.step(warm)
    b.sync()
    t0 = time.perf_counter()
    b.step(steps)
    b.sync()
    t_b = time.perf_counter() - t0
    out = {"cases": n_cases, "cells_per_step": cells, "steps": steps,
           "batched_cell_updates_per_s": cells * steps / t_b, "batched_ms_per_step": 1e3 * t_b / steps,
           "deck_write_s": t_write, "line_solve_rows_per_lane": tuned_rows}
    if oracle_check:
        k = min(range(n_cases), key=lambda i: members[i].cells)
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import _oracle as orc
        o = orc.OCase(dirs[k])
        o.step(warm + steps)
        ref = o.solution()
        got = b.solution(k)
        err = np.sqrt(((got - ref) ** 2).sum(0)) / np.maximum(np.sqrt((ref ** 2).sum(0)), 1e-300)
        out["oracle_check"] = {"case": members[k].name, "steps": warm + steps, "max_rel_l2_per_var": float(err.max())}
    b.close()
    # one case at a time through the single-case path (each step replayed from the case's own graph)
    t_s = 0.0
    for d in dirs:
        c = piafs_amd.Case(d)
        c.step(1)
        c.sync()
        t0 = time.perf_counter()
        c.step(serial_steps)
        c.sync()
        t_s += time.perf_counter() - t0
        c.close()
    out["serial_steps"] = serial_steps
    out["serial_cell_updates_per_s"] = cells * serial_steps / t_s
    out["speedup_vs_one_at_a_time"] = out["batched_cell_updates_per_s"] / out["serial_cell_updates_per_s"]
    return out


def bench_serve(tmp, n_requests=24, rate=None, rate_frac=0.7, batched_rate=None):
    """Poisson stream of solve requests (30% Sod 1D, N in [512, 8192], to
    t = 0.2; 70% 2D NS + photochemistry, side in [128, 1024], 200 steps)
    served with continuous batching under an HBM budget, against the same
    stream solved one request at a time in arrival order (FIFO)."""
    import threading
    import piafs_amd
    from piafs_amd import decks
    probe = decks.request_stream(n_requests, 1.0)
    work = sum(r.member.cells * r.member.steps for r in probe)
    if rate is None:
        # arrivals at rate_frac of the rate the batched engine can absorb
        rate = rate_frac * (batched_rate or 1.5e8) / (work / n_requests)
    reqs = decks.request_stream(n_requests, rate)
    dirs = _write_members(tmp, [r.member for r in reqs], "srv")
    lat = [None] * n_requests
    with piafs_amd.Server() as srv:
        ids = [None] * n_requests
        t_start = time.perf_counter()

        def feeder():
            for k, r in enumerate(reqs):
                dt = t_start + r.arrival - time.perf_counter()
                if dt > 0:
                    time.sleep(dt)
                ids[k] = srv.submit(dirs[k])

        th = threading.Thread(target=feeder)
        th.start()
        th.join()
        srv.wait_all()
        t_end = time.perf_counter()
        for k in range(n_requests):
            st, lat[k] = srv.poll(ids[k])
            if st != srv.DONE:
                raise RuntimeError(f"request {k} ({reqs[k].member.name}) ended in state {st}: {srv._err()}")
    lat = np.array(lat)
    out = {"requests": n_requests, "rate_per_s": rate, "cell_updates": work,
           "p50_latency_s": float(np.percentile(lat, 50)), "p99_latency_s": float(np.percentile(lat, 99)),
           "throughput_cell_updates_per_s": work / (t_end - t_start)}
    is1d = np.array([r.member.deck.solver["ndims"] == 1 for r in reqs])
    for tag, sel in (("1d", is1d), ("2d", ~is1d)):
        if sel.any():
            out[f"p50_latency_{tag}_s"] = float(np.percentile(lat[sel], 50))
            out[f"max_latency_{tag}_s"] = float(lat[sel].max())

    # the same stream, one request at a time in arrival order
    t_free, lat_s, busy, solo = 0.0, [], 0.0, []
    for k, r in enumerate(reqs):
        t0 = time.perf_counter()
        c = piafs_amd.Case(dirs[k])
        c.step(r.member.steps)
        c.sync()
        c.close()
        dur = time.perf_counter() - t0
        solo.append(dur)
        busy += dur
        start = max(r.arrival, t_free)
        t_free = start + dur
        lat_s.append(t_free - r.arrival)
    lat_s = np.array(lat_s)
    order = np.argsort(-lat)[:5]
    out["slowest_served"] = [{"request": reqs[k].member.name, "served_latency_s": round(float(lat[k]), 4),
                              "alone_s": round(solo[k], 4)} for k in order]
    out["serial_p50_latency_s"] = float(np.percentile(lat_s, 50))
    out["serial_p99_latency_s"] = float(np.percentile(lat_s, 99))
    out["serial_throughput_cell_updates_per_s"] = work / max(t_free, busy)
    out["p99_vs_serial"] = out["p99_latency_s"] / out["serial_p99_latency_s"]
    return out


SCHEME_VARIANTS = [  # (name, hyp_space_scheme, hyp_interp_type, upwinding)
    ("weno5_components_roe", "weno5", "components", "roe"),
    ("weno5_characteristic_rf", "weno5", "characteristic", "rf-char"),
    ("crweno5_components_llf", "crweno5", "components", "llf-char"),
    ("crweno5_characteristic_roe", "crweno5", "characteristic", "roe"),
    ("upw5_components_roe", "upw5", "components", "roe"),
]


def bench_schemes(tmp, nx=512, steps=20, warmup=2):
    """The interpolation / upwinding variants beyond the headline deck, on one
    moving-flow 2D deck with chemistry (periodic box, uniform flow (0.3, 0.2)
    plus a pressure bump: the reference's pivot-free block inverse of
    characteristic CRWENO5 needs a moving gas).  Throughput and the per-class
    kernel times of each variant."""
    import piafs_amd
    from piafs_amd import decks
    periodic = [("periodic", 0, 1, 0, 0, -1e99, 1e99), ("periodic", 0, -1, 0, 0, -1e99, 1e99),
                ("periodic", 1, 1, -1e99, 1e99, 0, 0), ("periodic", 1, -1, -1e99, 1e99, 0, 0)]
    out = {"grid": [nx, nx], "steps": steps, "deck": "periodic box, chemistry (nvars 11), Re 1000, RK4, "
           "flow (1, 0.3, 0.2) + 5 % pressure bump, dt 5e-4"}
    for name, scheme, interp, upw in SCHEME_VARIANTS:
        dk = decks.ns2d_open_box(nx=nx, ny=nx, n_iter=warmup + steps, dt=5e-4, flow=(1.0, 0.3, 0.2, None),
                                 scheme=scheme, interp=interp, upwinding=upw, include_chemistry=True, Re=1000.0,
                                 boundary=periodic)
        d = os.path.join(tmp, "scheme_" + name)
        dk.write(d)
        with piafs_amd.Case(d, device=0) as c:
            c.step(warmup)
            c.sync()
            c.profile(True)
            t0 = time.perf_counter()
            c.step(steps)
            c.sync()
            secs = time.perf_counter() - t0
            kt = c.kernel_times()
            c.profile(False)
            finite = bool(np.isfinite(c.solution()).all())
        out[name] = {"cell_updates_per_s": nx * nx * steps / secs, "ms_per_step": 1e3 * secs / steps,
                     "kernel_ms": kt, "finite": finite}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--nx", type=int, default=None,
                    help="grid side (same domain 10 pi, dx = 10 pi / nx; dt 0.01 * 1024 / nx)")
    ap.add_argument("--grid", type=int, default=1024, choices=(1024, 4096),
                    help="1024: 1024^2, dt 0.01 (default); 4096: 4096^2, dt 0.0025, 50 steps")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-profile", action="store_true", help="skip the per-class timing pass")
    ap.add_argument("--profile-steps", type=int, default=10,
                    help="steps of the per-class timing pass that follows the timed region")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default line-solve segment length (for profiling runs)")
    ap.add_argument("--no-batch", action="store_true", help="skip the ensemble leg")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving leg")
    ap.add_argument("--no-schemes", action="store_true", help="skip the interpolation / upwinding variants leg")
    ap.add_argument("--batch-cases", type=int, default=256)
    ap.add_argument("--batch-steps", type=int, default=20)
    ap.add_argument("--serve-requests", type=int, default=24)
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
        print("bench.py: piafs_amd runs on one GPU (nothing here uses a second GPU); "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import piafs_amd
    from piafs_amd import decks

    nx = args.nx or args.grid
    if args.grid == 4096 and args.steps == 200:
        args.steps = 50
    dt = 0.01 * 1024 / nx   # same acoustic CFL on the same domain
    with tempfile.TemporaryDirectory() as tmp:
        dk = decks.ns2d_photochem(nx=nx, n_iter=args.warmup + args.steps, dt=dt)
        d = os.path.join(tmp, "case")
        dk.write(d)
        case = piafs_amd.Case(d, device=0)
        inf = case.info
        nv, g = inf.nvars, inf.ghosts
        tuned_rows = None if args.no_autotune else case.autotune()[0]
        case.step(args.warmup)
        case.sync()
        # the timed region: steps replayed from the captured HIP graph, as a
        # caller of piafs_step runs them
        t0 = time.perf_counter()
        case.step(args.steps)
        case.sync()
        secs = time.perf_counter() - t0
        ktimes = {}
        if not args.no_profile:
            # per-class device times in a pass of their own right after it: the
            # same launches issued eagerly, HIP events around each on the stream
            # it runs on
            case.profile(True)
            case.step(args.profile_steps)
            case.sync()
            ktimes = case.kernel_times()
            case.profile(False)
        cells = nx * nx
        value = cells * args.steps / secs
        line = {
            "metric": METRIC, "value": value, "unit": "cell-updates/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": 1e3 * secs / args.steps, "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
            "config": {"workload": f"2D NavierStokes + photochemistry {nx}x{nx}, nvars {nv}, CRWENO5 components, "
                                   f"Roe, 4th-order viscous (Re 1000), RK4 dt {dt:g}",
                       "grid": [nx, nx], "nvars": nv, "steps": args.steps, "parallelism": "single GPU",
                       "line_solve_rows_per_lane": tuned_rows},
        }
        if ktimes:
            # per-class average launch time (ms) x launches per step -> share of the step
            ns = inf.nstages
            per_step = {"stage_bc": ns, "ghost": ns, "bflux": ns, "flux": ns, "recon_ends": ns, "recon": ns, "recon_full": ns, "recon_char": ns, "upwind": ns, "qderiv": ns,
                        "viscflux": ns, "assemble": ns, "update": 1}
            share = {k: v * per_step.get(k, 1) for k, v in ktimes.items()}
            dom = max(share.items(), key=lambda kv: kv[1])[0]
            paired = (not inf.characteristic) and dk.physics["upwinding"] in ("roe", "rusanov")
            io_byts = kernel_algorithmic_bytes(dom, nx, nx, nv, g, inf.nstages, True, paired)
            ms = ktimes[dom]
            # algorithmic bytes of one stage of a stage-fused RK step, 40 nv B per
            # cell (read the stage state and u_n, read and write the accumulator,
            # write the next stage state); the dominant kernel runs once per stage
            byts = STEP_BYTES_PER_CELL(nv, 1) * cells
            ach = byts / (ms * 1e-3) / 1e9
            line["roofline"] = {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                "frac": ach / HBM_PEAK_GBS,
                                "traffic": measured_traffic(nx, dom),
                                "kernel": dom, "kernel_ms": ms, "algorithmic_bytes_per_launch": byts,
                                "algorithmic_bytes_rule": "40 * nvars B per cell-update and stage",
                                "kernel_io_bytes_per_launch": io_byts,
                                "kernel_io_frac": (io_byts / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS) if io_byts else None}
            valu = valu_counters(nx, dom)
            if valu:
                line["roofline"].update(valu_fractions(valu, ms))
            line["kernel_ms"] = ktimes
            line["kernel_ms_per_step"] = share
        step_gbs = STEP_BYTES_PER_CELL(nv, inf.nstages) * value / 1e9
        line["step_roofline"] = {"bytes_per_cell_update": STEP_BYTES_PER_CELL(nv, inf.nstages),
                                 "achieved": step_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                 "frac": step_gbs / HBM_PEAK_GBS}
        if args.parity and nx == 1024 and args.warmup + args.steps == 200:
            fx = os.path.join(ROOT, "tests", "golden", "ns2d_chem_crweno_1024_200steps_ref.npz")
            if os.path.exists(fx):
                z = np.load(fx)
                U = case.solution().reshape(nx, nx, nv)
                sub = U[::32, ::32, :].reshape(-1, nv)
                ref = z["sub"].reshape(-1, nv)
                err = np.sqrt(((sub - ref) ** 2).sum(0)) / np.sqrt((ref ** 2).sum(0))
                line["parity"] = {"vs": "reference run (oracle/_ref) 1024^2 x 200 steps, stride-32 subsample",
                                  "max_rel_l2_per_var": float(err.max())}
        case.close()
        if not args.no_batch:
            line["batched"] = bench_batch(tmp, n_cases=args.batch_cases, steps=args.batch_steps)
        if not args.no_serve:
            br = line.get("batched", {}).get("batched_cell_updates_per_s")
            line["serve"] = bench_serve(tmp, n_requests=args.serve_requests, batched_rate=br)
        if not args.no_schemes:
            line["schemes"] = bench_schemes(tmp)
        if not args.no_cpu_baseline:
            line["cpu_baseline"] = cpu_baseline(tmp)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
