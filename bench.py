#!/usr/bin/env python3
"""Benchmark: particle-slice-steps/s of the fused GENESIS 1.3 v4 time-dependent
step on one MI355X.

Workload (time-dependent SASE, FFT field solver): 4096 slices x 16384
particles/slice, ngrid 256, gridmax 0.4 mm, Gaussian current (peak 4 kA,
sigma 1000 slices), 1e-4 linear energy chirp, shot noise (seed 1234569),
planar undulator lambdau 26 mm, aw 1.7 rms, gamma0 11350, FODO quadrupoles,
diagnostics recorded every step.  Synthetic, seeded input; state resident in
HBM before the timed region.

value  = mixed-precision throughput (FP32 storage/push, FP64 phase, Thomas and
         reductions); the FP64 path is timed on the same workload and reported
         beside it with the mixed-vs-FP64 slice power / bunching error.
A step = one lattice step of the whole beam: particle pass (transverse half
steps, gather + RK4, deposition, beam diagnostics), FFT row and column
passes (field diagnostics fused), slippage.

    python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "particle-slice-steps/s, mixed vs FP64 at ≤1e-3 power err; p50/p99 req latency"
PEAK_HBM_GBS = 8000.0

LAMBDAU, AW, GAMMA0 = 0.026, 1.7, 11350.0
LAMBDA0 = LAMBDAU * (1 + AW * AW) / (2 * GAMMA0 * GAMMA0)


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def workload(nslice: int, npart: int):
    from genesis4_cuda_port_amd import api
    s = np.arange(nslice, dtype=np.float64)
    current = 4000.0 * np.exp(-0.5 * ((s - 0.5 * (nslice - 1)) / 1000.0) ** 2)
    beam = api.load_beam(nslice=nslice, npart=npart, nbins=4, gamma0=GAMMA0, espread_rel=1e-4,
                         emitn=0.4e-6, beta=15.0, slicelength=LAMBDA0, current=current,
                         shotnoise=True, seed=1234569, chirp_rel=1e-4)
    steps = api.lattice_fodo(LAMBDAU, [AW] * 10, 100, 46, 3, 1.5, 1)
    return beam, steps


def algorithmic_bytes(cls: str, prec: int, np_: int, cells: int) -> float:
    """Compulsory HBM bytes per launch of each kernel class.  Particle pass
    (transverse + push + deposition + beam sums, fused): particle state read
    and written, 2 (5 s_p + s_theta) per particle, plus the source plane
    written once, s_c per cell; field gathers excluded.  FFT row pass: field
    read + written; column pass: field read + written and the source read.
    The kernels touch only the deposition footprint of the source planes, so
    their measured traffic can be below these figures."""
    sp = 8 if prec == 0 else 4     # particle scalar
    sc = 16 if prec == 0 else 8    # complex cell
    if cls == "particle":
        return np_ * 2.0 * (5 * sp + 8) + cells * sc
    if cls == "fft_rows":
        return 2.0 * cells * sc
    if cls in ("fft_cols", "adi_x", "adi_y"):
        return 3.0 * cells * sc
    return 0.0


TRAFFIC_PROFILE = ROOT / "profiles" / "hbm_traffic_fft256_4096x16384_mixed.md"
TRAFFIC_KERNEL = {"particle": "particle_kernel<float, 1, true, 1, false>",
                  "fft_rows": "fft_rows_reg_kernel<float, 256>",
                  "fft_cols": "fft_cols_reg_kernel<float, 256>"}


def measured_traffic(cls: str):
    """HBM bytes per launch of a MIXED kernel class at the benchmark workload,
    from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary (already
    scaled by the copy-kernel calibration); None when not profiled."""
    name = TRAFFIC_KERNEL.get(cls)
    if name is None or not TRAFFIC_PROFILE.exists():
        return None
    for line in TRAFFIC_PROFILE.read_text().splitlines():
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cells) == 5 and cells[0] == f"`{name}`":
            return float(cells[4]) * 1e9
    return None


def run_gpu(prec: int, beam, steps, ngrid: int, gridmax: float, warmup: int, nsteps: int,
            solver=None):
    from genesis4_cuda_port_amd import api
    from genesis4_cuda_port_amd._abi import KERNEL_CLASSES
    nslice = beam.offsets.size - 1
    with api.Context(0) as ctx:
        b = api.Beam(ctx, prec, beam.offsets, beam.current, GAMMA0)
        b.upload(beam.x, beam.y, beam.px, beam.py, beam.gamma, beam.theta)
        f = api.Field(ctx, prec, nslice, ngrid, gridmax, LAMBDA0, LAMBDA0)
        f.init_solver(api.FFT if solver is None else solver)
        nrun = warmup + nsteps
        run = api.Run(ctx, b, f, list(steps)[:nrun], LAMBDA0, 1.0, True, False, 1)
        run.advance(warmup)
        ctx.synchronize()
        ctx.profile(True)
        ctx.synchronize()
        t0 = time.perf_counter()
        run.advance(nsteps)
        ctx.synchronize()
        t1 = time.perf_counter()
        kern = {}
        for k, name in enumerate(KERNEL_CLASSES):
            n, ms = ctx.profile_read(k)
            if n:
                kern[name] = {"launches": n, "total_ms": ms, "avg_ms": ms / n}
        ctx.profile(False)
        braw, fraw = run.download_diag()
        mem = ctx.memory()
        run.close()
        f.close()
        b.close()
    return (t1 - t0), kern, braw, fraw, mem


def power_bunching(braw, fraw, current, ngrid, gridmax):
    from genesis4_cuda_port_amd import api
    dgrid = 2 * gridmax / (ngrid - 1)
    fq, _ = api.finalize_field(fraw, dgrid, LAMBDA0, LAMBDA0)
    bq, _ = api.finalize_beam(braw, 1, current, GAMMA0)
    return fq[:, 0], bq[:, 25]


def cpu_baseline(beam, steps, ngrid, gridmax, nslice_cpu: int, nsteps_cpu: int,
                 threads: int = 0, keep: bool = False):
    """The FP64 oracle on the host cores over the first nslice_cpu slices of the
    same workload (one untimed step, then nsteps_cpu timed steps).  With keep,
    also returns the oracle's last beam and field records (after
    1 + nsteps_cpu steps): slippage carries the field from the tail to the
    head, so the first slices of the full beam evolve as this sub-beam does,
    and the GPU runs' records of those slices compare with these."""
    sys.path.insert(0, str(ROOT / "tests"))
    import _oracle
    # every core this process may use: the CPU share the host gives the job
    # (OMP_NUM_THREADS where the host sets it, else the affinity mask)
    avail = len(os.sched_getaffinity(0))
    if threads <= 0:
        threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or avail
    _oracle.lib().orc_set_threads(threads)
    sub = beam.copy()
    npart = int(beam.offsets[1] - beam.offsets[0])
    n = nslice_cpu * npart
    for k in ("x", "y", "px", "py", "gamma", "theta"):
        setattr(sub, k, getattr(beam, k)[:n].copy())
    sub.offsets = beam.offsets[:nslice_cpu + 1].copy()
    sub.current = beam.current[:nslice_cpu].copy()
    of = _oracle.OField(nslice_cpu, ngrid, gridmax, LAMBDA0, LAMBDA0)
    st = list(steps)
    from genesis4_cuda_port_amd.api import steps_array
    warm = steps_array(st[:1])
    timed = steps_array(st[1:1 + nsteps_cpu])
    sub, _, _ = _oracle.run(sub, of, warm, gamma0=GAMMA0, lambda0=LAMBDA0, solver=1)
    t0 = time.perf_counter()
    _, braw, fraw = _oracle.run(sub, of, timed, gamma0=GAMMA0, lambda0=LAMBDA0, solver=1)
    dt = time.perf_counter() - t0
    res = {"value": n * nsteps_cpu / dt, "unit": "particle-slice-steps/s", "cores": threads,
           "kind": "port", "host_nproc": os.cpu_count(), "affinity_cpus": avail,
           "omp_num_threads_env": os.environ.get("OMP_NUM_THREADS"),
           "sample": f"FP64 oracle, first {nslice_cpu} of 4096 slices x {npart} particles, "
                     f"ngrid {ngrid} FFT, {nsteps_cpu} steps after 1 untimed ({dt:.1f} s)"}
    return (res, braw[-1], fraw[-1]) if keep else res


def slice_quantities(braw, fraw, current, ngrid, gridmax):
    """Per-slice power, bunching and energy spread of one record."""
    from genesis4_cuda_port_amd import api
    dgrid = 2 * gridmax / (ngrid - 1)
    fq, _ = api.finalize_field(np.ravel(fraw), dgrid, LAMBDA0, LAMBDA0)
    bq, _ = api.finalize_beam(np.ravel(braw), 1, current, GAMMA0)
    return {"power": fq[:, 0], "bunching": bq[:, 25], "energyspread": bq[:, 1]}


def rel_errors(got, ref):
    return {f"{k}_rel": float(np.linalg.norm(got[k] - ref[k]) / max(np.linalg.norm(ref[k]), 1e-300))
            for k in ref}


def _run_records(prec, beam, steps, ngrid, gridmax, re, im, solver, timerun):
    """Records of a whole run on the GPU (every step an output step)."""
    from genesis4_cuda_port_amd import api
    S = beam.offsets.size - 1
    with api.Context(0) as ctx:
        b = api.Beam(ctx, prec, beam.offsets, beam.current, GAMMA0)
        b.upload(beam.x, beam.y, beam.px, beam.py, beam.gamma, beam.theta)
        f = api.Field(ctx, prec, S, ngrid, gridmax, LAMBDA0, LAMBDA0)
        f.upload(re, im, 0)
        f.init_solver(solver)
        run = api.Run(ctx, b, f, steps, LAMBDA0, 1.0, timerun, False, 1)
        run.advance(len(steps))
        braw, fraw = run.download_diag()
        run.close()
        f.close()
        b.close()
    return braw, fraw


def precision_along_z(nslice_td: int = 512):
    """MIXED against FP64 over the whole 10 x 100-period lattice: per record,
    the normwise relative difference of slice power, bunching and energy
    spread, for (a) the steady-state run of 1 slice x 8192 particles, ngrid
    151, ADI, 1 kW seed, and (b) the time-dependent SASE workload of the
    benchmark (Gaussian current, chirp, shot noise, ngrid 256, FFT) on a
    window of nslice_td slices at the bunch centre.  Reports the errors every
    50 steps and the largest z up to which all three stay within 1e-3
    (records after the first step; bunching relative to max(|b|, 1/sqrt(Ne)),
    the shot-noise level)."""
    from genesis4_cuda_port_amd import api
    steps = api.lattice_fodo(LAMBDAU, [AW] * 10, 100, 46, 3, 1.5, 1)
    dz = LAMBDAU
    out = {}
    cases = []
    b0 = api.load_beam(nslice=1, npart=8192, nbins=4, gamma0=GAMMA0, espread_rel=1e-4,
                       emitn=0.4e-6, beta=15.0, slicelength=LAMBDA0, current=3000.0,
                       shotnoise=False, seed=1234567)
    re0, im0 = api.seed_gaussian(1, 151, 0.3e-3, LAMBDA0, 1e3, 50e-6)
    cases.append(("steady_state_1x8192_adi151", b0, 151, 0.3e-3, re0, im0, api.ADI, False))
    S = nslice_td
    s = np.arange(S, dtype=np.float64) + 2048 - S / 2
    cur = 4000.0 * np.exp(-0.5 * ((s - 2047.5) / 1000.0) ** 2)
    b2 = api.load_beam(nslice=S, npart=16384, nbins=4, gamma0=GAMMA0, espread_rel=1e-4,
                       emitn=0.4e-6, beta=15.0, slicelength=LAMBDA0, current=cur, shotnoise=True,
                       seed=1234569, chirp_rel=1e-4 * S / 4096)
    z2 = np.zeros(S * 256 * 256)
    cases.append((f"sase_{S}x16384_fft256", b2, 256, 0.4e-3, z2, z2, api.FFT, True))
    for name, beam, G, gm, re, im, solver, td in cases:
        q = {}
        for pname, prec in (("fp64", api.FP64), ("mixed", api.MIXED)):
            braw, fraw = _run_records(prec, beam, steps, G, gm, re, im, solver, td)
            rows = [slice_quantities(braw[k], fraw[k], beam.current, G, gm)
                    for k in range(braw.shape[0])]
            q[pname] = rows
        # Bunching below the shot-noise level 1/sqrt(Ne) of a slice (Ne
        # electrons) carries no physical information (a mirrored quiet start
        # has a bunching of round-off size until the field has acted), so
        # its error is taken relative to max(|b|, 1/sqrt(Ne)).
        ne = float(np.mean(beam.current)) * LAMBDA0 / 1.602176634e-19 / 2.99792458e8
        floor = {"power": 1e-300, "bunching": np.sqrt(beam.current.size / ne),
                 "energyspread": 1e-300}
        errs = {k: [] for k in ("power", "bunching", "energyspread")}
        for a, b in zip(q["mixed"], q["fp64"]):
            for k in errs:
                den = max(np.linalg.norm(b[k]), floor[k])
                errs[k].append(float(np.linalg.norm(a[k] - b[k]) / den))
        n = len(errs["power"])
        worst = [max(errs[k][r] for k in errs) for r in range(n)]
        bad = [r for r in range(1, n) if worst[r] > 1e-3]
        z_ok = (bad[0] - 1) * dz if bad else (n - 1) * dz
        upto = worst[1:bad[0]] if bad else worst[1:]
        p64 = [float(np.sum(r["power"])) for r in q["fp64"]]
        out[name] = {"z_m": [round(r * dz, 4) for r in range(0, n, 50)],
                     "power_W_fp64": [p64[r] for r in range(0, n, 50)],
                     **{f"{k}_rel": [errs[k][r] for r in range(0, n, 50)] for k in errs},
                     "z_within_1e-3_m": z_ok, "lattice_m": (n - 1) * dz,
                     "max_rel_up_to_z": max(upto) if upto else None}
    return out


def time_dependent_adi(nsteps: int, warmup: int):
    """Time-dependent SASE in FP64 with the ADI solver: 1024 slices x 8192
    particles/slice, ngrid 151, gridmax 0.3 mm, flat-top 3 kA, shot noise
    (seed 1234568), same undulator and lattice as the main workload."""
    from genesis4_cuda_port_amd import api
    nslice, npart, ngrid, gm = 1024, 8192, 151, 0.3e-3
    beam = api.load_beam(nslice=nslice, npart=npart, nbins=4, gamma0=GAMMA0, espread_rel=1e-4,
                         emitn=0.4e-6, beta=15.0, slicelength=LAMBDA0, current=3000.0,
                         shotnoise=True, seed=1234568)
    steps = api.lattice_fodo(LAMBDAU, [AW] * 10, 100, 46, 3, 1.5, 1)
    t, kern, _, _, _ = run_gpu(api.FP64, beam, steps, ngrid, gm, warmup, nsteps, solver=api.ADI)
    np_ = nslice * npart
    cells = nslice * ngrid * ngrid
    return {"workload": f"time-dependent SASE, ADI, FP64, {nslice} slices x {npart} "
                        f"particles/slice, ngrid {ngrid}",
            "value": np_ * nsteps / t, "unit": "particle-slice-steps/s",
            "ms_per_step": t / nsteps * 1e3,
            "kernels": {k: dict(v, gbs=algorithmic_bytes(k, api.FP64, np_, cells) /
                                (v["avg_ms"] * 1e-3) / 1e9) for k, v in kern.items()}}


def taper_scan(prec: int, nruns: int, seed: int):
    """Batched taper scan: nruns independent steady-state runs (grid drawn
    from 101/151/201/255, 4096/8192/16384 particles, aw(z) = aw0 (1 - a z -
    b z^2) per module with a, b drawn from [0, 2e-3]) over 10 modules of 100
    periods with FODO focusing, ADI, MIXED or FP64, records every 10 steps.
    Returns throughput in particle-slice-steps/s of the advance (state and
    tables already on the device)."""
    from genesis4_cuda_port_amd import api
    rng = np.random.default_rng(seed)
    nmod = 10
    steps = api.lattice_fodo(LAMBDAU, [AW] * nmod, 100, 46, 3, 1.5, 10)
    ns = len(steps)
    # module index of each step (100 one-period steps per module)
    z = np.floor(np.arange(ns) * nmod / ns) / nmod
    runs = []
    for _ in range(nruns):
        a, b = rng.uniform(0.0, 2e-3, 2)
        runs.append(dict(npart=int(rng.choice([4096, 8192, 16384])), nbins=4,
                         ngrid=int(rng.choice([101, 151, 201, 255])), gridmax=0.3e-3,
                         current=3000.0, gamma0=GAMMA0, espread_rel=1e-4, emitn=0.4e-6,
                         beta=15.0, shotnoise=False, seed=int(rng.integers(1, 1 << 31)),
                         seed_power=1e3, seed_w0=50e-6, aw=AW * (1 - a * z - b * z * z)))
    with api.Context(0) as ctx:
        batch = api.Batch(ctx, prec, api.ADI, runs, steps, LAMBDA0)
        batch.advance(1)
        ctx.synchronize()
        t = time.perf_counter()
        batch.advance(ns - 1)
        ctx.synchronize()
        dt = time.perf_counter() - t
        batch.close()
    parts = sum(r["npart"] for r in runs)
    grids = sorted({r["ngrid"] for r in runs})
    return {"value": parts * (ns - 1) / dt, "unit": "particle-slice-steps/s",
            "ms_per_step": dt / (ns - 1) * 1e3, "runs": nruns, "steps": ns - 1,
            "particles": parts, "grids": grids}


def serving_requests(n: int, seed: int):
    """The request stream: 70 % steady-state runs (grid 101/151/201/255,
    4096/8192/16384 particles, linear + quadratic aw taper with coefficients
    drawn from [0, 2e-3]) and 30 % time-dependent runs (128-2048 slices,
    grid 101-255, 4096-16384 particles/slice), all on one 100-period
    undulator module with FODO focusing, ADI field solver, records every 10
    steps."""
    from genesis4_cuda_port_amd import api
    rng = np.random.default_rng(seed)
    steps = api.lattice_fodo(LAMBDAU, [AW], 100, 46, 3, 1.5, 10)
    z = np.arange(1, len(steps) + 1) / len(steps)
    out = []
    for _ in range(n):
        if rng.random() < 0.7:
            a, b = rng.uniform(0.0, 2e-3, 2)
            out.append(dict(kind=0, nslice=1, ngrid=int(rng.choice([101, 151, 201, 255])),
                            npart=int(rng.choice([4096, 8192, 16384])), aw=AW * (1 - a * z - b * z * z),
                            seed=int(rng.integers(1, 1 << 31)), seed_power=1e3))
        else:
            out.append(dict(kind=1, nslice=int(rng.integers(128, 2049)),
                            ngrid=int(rng.integers(101, 256)),
                            npart=4 * int(rng.integers(1024, 4097)),
                            seed=int(rng.integers(1, 1 << 31)), seed_power=0.0))
    return steps, out


def serving(prec: int, n: int, seed: int, rate_hz):
    """Submit the stream to a request loop (200 GB HBM budget, 4 stream
    lanes): all at once when rate_hz is None (capacity), else with Poisson
    arrivals at rate_hz.  Returns the server's latency / throughput stats."""
    from genesis4_cuda_port_amd import api
    steps, reqs = serving_requests(n, seed)
    srv = api.Server(0, hbm_budget=200 * 10 ** 9, nstreams=4)
    rng = np.random.default_rng(seed + 1)
    t0 = time.perf_counter()
    t_next = 0.0
    ids = []
    for r in reqs:
        if rate_hz:
            t_next += rng.exponential(1.0 / rate_hz)
            dt = t0 + t_next - time.perf_counter()
            if dt > 0:
                time.sleep(dt)
        ids.append(srv.submit(precision=prec, solver=api.ADI, steps=steps, gridmax=0.3e-3,
                              current=3000.0, gamma0=GAMMA0, lambda0=LAMBDA0, espread_rel=1e-4,
                              **r))
    for i in ids:
        srv.wait(i)
    wall = time.perf_counter() - t0
    st = srv.stats()
    srv.close()
    return {"p50_ms": st["p50_s"] * 1e3, "p99_ms": st["p99_s"] * 1e3,
            "p99_over_n": f"p99 of {int(st['completed'])} requests",
            "mean_ms": st["mean_s"] * 1e3, "requests": int(st["completed"]),
            "particle_slice_steps_per_s": st["particle_slice_steps_per_s"],
            "requests_per_s": len(reqs) / wall}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--slices", type=int, default=4096)
    ap.add_argument("--particles", type=int, default=16384)
    ap.add_argument("--ngrid", type=int, default=256)
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--serve-requests", type=int, default=200)
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-adi", action="store_true")
    ap.add_argument("--no-precision", action="store_true",
                    help="skip the MIXED-vs-FP64 comparison over the whole lattice")
    ap.add_argument("--batch-runs", type=int, default=96)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-slices", type=int, default=1536)
    ap.add_argument("--cpu-steps", type=int, default=12)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run 3 plain field copies of 4096 x 256^2 FP32 cells (2 GiB "
                         "read + 2 GiB written each) to calibrate the HBM byte counters")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {a.gpus}); this build has no "
              "multi-GPU path.  The slice axis would shard across GPUs with one plane handed "
              "over per slippage event; see the design notes.", file=sys.stderr)
        return 2
    from genesis4_cuda_port_amd import api
    if api.device_count() < 1:
        print("bench.py: no GPU visible to libg4hip", file=sys.stderr)
        return 3
    gridmax = 0.4e-3
    if a.copy_calibration:
        with api.Context(0) as ctx:
            src = api.Field(ctx, api.MIXED, a.slices, a.ngrid, gridmax, LAMBDA0, LAMBDA0)
            dst = api.Field(ctx, api.MIXED, a.slices, a.ngrid, gridmax, LAMBDA0, LAMBDA0)
            for _ in range(3):
                dst.copy_from(src)
            ctx.synchronize()
            dst.close()
            src.close()
            nbytes = a.slices * a.ngrid * a.ngrid * 8
            print(json.dumps({"copy_bytes_read": nbytes, "copy_bytes_written": nbytes,
                              "copies": 3}), flush=True)
        return 0
    log(f"loading beam {a.slices} x {a.particles}")
    t = time.perf_counter()
    beam, steps = workload(a.slices, a.particles)
    log(f"beam loaded in {time.perf_counter() - t:.1f} s")
    np_total = a.slices * a.particles
    cells = a.slices * a.ngrid * a.ngrid

    log("mixed precision run")
    tm, km, bm, fm, mem_m = run_gpu(api.MIXED, beam, steps, a.ngrid, gridmax, a.warmup, a.steps)
    value = np_total * a.steps / tm
    log(f"mixed: {tm / a.steps * 1e3:.3f} ms/step, {value:.3e} p-s-s/s")
    out = {
        "metric": METRIC, "value": value, "unit": "particle-slice-steps/s", "n_gpus": 1,
        "steps": a.steps, "warmup": a.warmup, "ms_per_step": tm / a.steps * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32+f64 (mixed: f32 state/field/push, f64 phase/Thomas/reductions)",
        "data": "synthetic (seeded quiet-start beam with shot noise)",
        "config": {"workload": "time-dependent SASE, FFT field solver, "
                               f"{a.slices} slices x {a.particles} particles/slice, "
                               f"ngrid {a.ngrid}, mixed precision",
                   "nslice": a.slices, "npart_per_slice": a.particles, "ngrid": a.ngrid,
                   "solver": "fft", "diagnostics": "every step", "parallelism": "single GPU"},
    }
    # roofline of the dominant kernel class of the mixed run
    dom = max(km.items(), key=lambda kv: kv[1]["total_ms"])
    name, st = dom
    ab = algorithmic_bytes(name, api.MIXED, np_total, cells)
    achieved = ab / (st["avg_ms"] * 1e-3) / 1e9
    out["roofline"] = {"bound": "hbm", "kernel": name, "achieved": achieved,
                       "peak": PEAK_HBM_GBS, "unit": "GB/s", "frac": achieved / PEAK_HBM_GBS,
                       "traffic": measured_traffic(name), "algorithmic_bytes_per_launch": ab,
                       "avg_launch_ms": st["avg_ms"]}
    # whole step against the same model (particle + row + column pass)
    sb = sum(algorithmic_bytes(k, api.MIXED, np_total, cells)
             for k in ("particle", "fft_rows", "fft_cols"))
    out["step_roofline"] = {"mixed": {"bytes_per_step": sb,
                                      "achieved": sb / (tm / a.steps) / 1e9,
                                      "frac": sb / (tm / a.steps) / 1e9 / PEAK_HBM_GBS}}
    out["kernels_mixed"] = {k: dict(v, gbs=algorithmic_bytes(k, api.MIXED, np_total, cells) /
                                    (v["avg_ms"] * 1e-3) / 1e9) for k, v in km.items()}
    out["hbm_bytes_resident_mixed"] = mem_m

    if not a.no_fp64:
        log("FP64 run")
        t64, k64, b64, f64, mem_d = run_gpu(api.FP64, beam, steps, a.ngrid, gridmax, a.warmup,
                                            a.steps)
        v64 = np_total * a.steps / t64
        log(f"fp64: {t64 / a.steps * 1e3:.3f} ms/step, {v64:.3e} p-s-s/s")
        pm, bmq = power_bunching(bm[-1].ravel(), fm[-1].ravel(), beam.current, a.ngrid, gridmax)
        pd, bdq = power_bunching(b64[-1].ravel(), f64[-1].ravel(), beam.current, a.ngrid, gridmax)
        out["fp64"] = {"value": v64, "ms_per_step": t64 / a.steps * 1e3,
                       "kernels": {k: dict(v, gbs=algorithmic_bytes(k, api.FP64, np_total, cells) /
                                           (v["avg_ms"] * 1e-3) / 1e9) for k, v in k64.items()}}
        sb64 = sum(algorithmic_bytes(k, api.FP64, np_total, cells)
                   for k in ("particle", "fft_rows", "fft_cols"))
        out["step_roofline"]["fp64"] = {"bytes_per_step": sb64,
                                        "achieved": sb64 / (t64 / a.steps) / 1e9,
                                        "frac": sb64 / (t64 / a.steps) / 1e9 / PEAK_HBM_GBS}
        out["mixed_over_fp64"] = value / v64
        out["mixed_vs_fp64_error"] = {
            "power_rel": float(np.linalg.norm(pm - pd) / max(np.linalg.norm(pd), 1e-300)),
            "bunching_rel": float(np.linalg.norm(bmq - bdq) / max(np.linalg.norm(bdq), 1e-300)),
            "at_step": a.warmup + a.steps}
    if not a.no_cpu:
        log("cpu baseline (FP64 oracle)")
        res, ob, of_ = cpu_baseline(beam, steps, a.ngrid, gridmax, a.cpu_slices, a.cpu_steps,
                                    keep=True)
        out["cpu_baseline"] = res
        # the GPU runs against the FP64 CPU oracle on the same slices and step
        n, k = a.cpu_slices, 1 + a.cpu_steps
        if k < len(bm):
            cur = beam.current[:n]
            ref = slice_quantities(ob, of_, cur, a.ngrid, gridmax)
            vs = {"slices": n, "at_step": k,
                  "mixed": rel_errors(slice_quantities(bm[k][:n], fm[k][:n], cur, a.ngrid,
                                                       gridmax), ref)}
            if not a.no_fp64:
                vs["fp64"] = rel_errors(slice_quantities(b64[k][:n], f64[k][:n], cur, a.ngrid,
                                                         gridmax), ref)
            out["vs_cpu_oracle"] = vs
        # and one host thread on a smaller sample (the scalar oracle)
        one = cpu_baseline(beam, steps, a.ngrid, gridmax, max(1, a.cpu_slices // 32),
                           a.cpu_steps, threads=1)
        out["cpu_baseline"]["single_thread"] = {k: one[k] for k in ("value", "cores", "sample")}
    if not a.no_precision:
        log("MIXED vs FP64 along the whole lattice")
        out["mixed_vs_fp64_along_z"] = precision_along_z()
    if not a.no_adi:
        log("time-dependent ADI run (FP64)")
        out["time_dependent_adi_fp64"] = time_dependent_adi(a.steps, a.warmup)
    if not a.no_batch:
        scan = {"workload": f"taper scan of {a.batch_runs} steady-state runs in one batch "
                            "(uneven grid and particle count, 10 x 100-period modules, ADI)",
                "seed": 1234573}
        for name, prec in (("mixed", api.MIXED), ("fp64", api.FP64)):
            log(f"taper scan ({name})")
            scan[name] = taper_scan(prec, a.batch_runs, 1234573)
        out["taper_scan"] = scan
    out["latency_p50_p99"] = None
    if not a.no_serve:
        # capacity with every request queued at once, then Poisson arrivals at
        # 80 % of that capacity; MIXED and FP64
        serve = {"workload": "request stream of serving_requests(): 70 % steady-state "
                             "(batched), 30 % time-dependent; 100-step lattice; ADI; "
                             "200 GB HBM budget, 4 stream lanes", "seed": 1234571}
        for name, prec in (("mixed", api.MIXED), ("fp64", api.FP64)):
            log(f"serving capacity ({name})")
            cap = serving(prec, a.serve_requests, 1234571, None)
            rate = 0.8 * cap["requests_per_s"]
            log(f"serving at {rate:.2f} req/s ({name})")
            res = serving(prec, a.serve_requests, 1234571, rate)
            res["rate_hz"] = rate
            res["capacity_requests_per_s"] = cap["requests_per_s"]
            serve[name] = res
        out["latency_p50_p99"] = serve
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
