"""GPU parity at the shapes of the five workload configurations.

Each test runs the HIP path through the C ABI at a configuration's own
grid, particle count, solver and lattice (with fewer slices or steps where
the FP64 oracle would otherwise take minutes) and compares it with the FP64
CPU oracle on the same seeded inputs:

  0. steady state, 1 slice x 8192 particles, ngrid 151, ADI, FP64, 1 kW
     seed, the whole 10 x 100-period FODO lattice: slice power and
     bunching within 1e-10 on every record;
  1. time-dependent SASE, 32 slices x 8192, ngrid 151, ADI, FP64, shot
     noise, 60 steps: cell indices, slippage shifts, ring offset and
     particle counts bit-exact, power and bunching within 1e-10;
  2. time-dependent SASE, 32 slices x 16384, ngrid 256, gridmax 0.4 mm,
     FFT, Gaussian current and 1e-4 chirp, 40 steps through g4h_run: FP64
     within 1e-10, MIXED within 1e-3;
  3. one batch of 8 tapered steady-state runs spanning ngrid 101-255 and
     4096-16384 particles over the 10-module lattice: every run against
     the oracle, FP64 within 1e-10, MIXED within 1e-3;
  4. a seeded Poisson stream of 16 requests (70 % steady state, 30 %
     time-dependent with 128-512 slices) through the request loop: each
     result equal to the request run alone (FP64 1e-12, MIXED 1e-6) and one
     request of each kind against the oracle (FP64 1e-10, MIXED 1e-3).

Errors are normwise relative over slices; "every record" means the bound
holds for each record separately.  A mirrored quiet start has a bunching of
round-off size until the field has acted, so the bunching of a record is
compared against max(|bunching|, floor): 1e-6 in FP64, and in MIXED the
shot-noise level 1/sqrt(Ne) of the slice's Ne electrons (below which a
bunching value carries no physical information).
"""
from __future__ import annotations

import concurrent.futures as cf
import time

import numpy as np
import pytest

import _oracle
from _oracle import OField, dp, ip
from cases import AW, BETA, EMITN, ESPREAD, GAMMA0, LAMBDA0, LAMBDAU
from genesis4_cuda_port_amd import api

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-10
MIXED_TOL = 1e-3
GM151 = 0.3e-3
GM256 = 0.4e-3
E_CHARGE, C_LIGHT = 1.602176634e-19, 2.99792458e8


def shot_noise_bunching(current, slicelength=LAMBDA0):
    """1/sqrt(Ne) of a slice of `current` amperes and `slicelength` metres."""
    return 1.0 / np.sqrt(current * slicelength / E_CHARGE / C_LIGHT)


def relerr(a, b):
    a = np.ravel(np.asarray(a, dtype=np.float64))
    b = np.ravel(np.asarray(b, dtype=np.float64))
    den = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (den if den > 0 else 1.0)


def fodo(nmod, output_step=1, aw=None):
    """The configurations' lattice: nmod modules of 100 one-period steps,
    FODO quadrupoles every 46 periods (1.2 m), 3 periods long, k1 1.5."""
    return api.lattice_fodo(LAMBDAU, [AW] * nmod if aw is None else aw, 100, 46, 3, 1.5,
                            output_step)


def records(braw, fraw, current, ngrid, gridmax, nharm=1):
    """Per-record slice power [W] and bunching |<exp(i theta)>| (arrays
    [records, slices])."""
    dg = 2 * gridmax / (ngrid - 1)
    n, S = braw.shape[0], braw.shape[1]
    P = np.empty((n, S))
    B = np.empty((n, S))
    for r in range(n):
        fq, _ = api.finalize_field(np.ravel(fraw[r]), dg, LAMBDA0, LAMBDA0)
        bq, _ = api.finalize_beam(np.ravel(braw[r]), nharm, current, GAMMA0)
        P[r] = fq[:, 0]
        B[r] = bq[:, 25]
    return P, B


def per_record(got, ref, floor=0.0):
    """Largest per-record normwise relative error."""
    worst = 0.0
    for g, r in zip(got, ref):
        den = max(np.linalg.norm(r), floor * np.sqrt(r.size))
        worst = max(worst, np.linalg.norm(g - r) / (den if den > 0 else 1.0))
    return worst


def gpu_run(ctx, prec, solver, beam, ngrid, gridmax, steps, re, im, *, timerun, sample=1.0,
            keep=False):
    S = beam.offsets.size - 1
    b = api.Beam(ctx, prec, beam.offsets, beam.current, GAMMA0)
    b.upload(beam.x, beam.y, beam.px, beam.py, beam.gamma, beam.theta)
    f = api.Field(ctx, prec, S, ngrid, gridmax, LAMBDA0, LAMBDA0 * sample)
    f.upload(re, im, 0)
    f.init_solver(solver)
    run = api.Run(ctx, b, f, steps, LAMBDA0, sample, timerun, False, 1)
    assert run.advance(len(steps)) == len(steps)
    braw, fraw = run.download_diag()
    run.close()
    if keep:
        return braw, fraw, b, f
    f.close()
    b.close()
    return braw, fraw


def oracle_run(beam, ngrid, gridmax, steps, re, im, *, solver, timerun, sample=1.0):
    S = beam.offsets.size - 1
    of = OField(S, ngrid, gridmax, LAMBDA0, LAMBDA0 * sample, 1, re, im)
    ob, obraw, ofraw = _oracle.run(beam, of, steps, gamma0=GAMMA0, lambda0=LAMBDA0,
                                   solver=solver, sample=sample, timerun=timerun)
    return ob, of, obraw, ofraw


# ---- configuration 0 ---------------------------------------------------------

def test_config0_steady_state_full_lattice(gpu_ctx):
    print("case config0: 1 x 8192, ngrid 151, ADI, FP64, 1000 steps", flush=True)
    steps = api.steps_array(list(fodo(10)))
    assert len(steps) == 1000
    beam = api.load_beam(nslice=1, npart=8192, nbins=4, gamma0=GAMMA0, espread_rel=ESPREAD,
                         emitn=EMITN, beta=BETA, slicelength=LAMBDA0, current=3000.0,
                         shotnoise=False, seed=1234567)
    re, im = api.seed_gaussian(1, 151, GM151, LAMBDA0, 1e3, 50e-6)
    t = time.perf_counter()
    braw, fraw = gpu_run(gpu_ctx, api.FP64, api.ADI, beam, 151, GM151, steps, re, im,
                         timerun=False)
    t_gpu = time.perf_counter() - t
    t = time.perf_counter()
    _, _, obraw, ofraw = oracle_run(beam, 151, GM151, steps, re, im, solver=api.ADI,
                                    timerun=False)
    t_orc = time.perf_counter() - t
    assert braw.shape == obraw.shape == (1001, 1, braw.shape[2])
    P, B = records(braw, fraw, beam.current, 151, GM151)
    oP, oB = records(obraw, ofraw, beam.current, 151, GM151)
    ep, eb = per_record(P, oP), per_record(B, oB, floor=1e-6)
    print(f"  gpu {t_gpu:.1f} s, oracle {t_orc:.1f} s; power gain {oP[-1, 0] / oP[0, 0]:.3e}, "
          f"final bunching {oB[-1, 0]:.3f}; worst record: power {ep:.2e} bunching {eb:.2e}",
          flush=True)
    assert ep < FP64_TOL
    assert eb < FP64_TOL


# ---- configuration 1 ---------------------------------------------------------

def test_config1_time_dependent_adi(gpu_ctx):
    S, P_, G, NST = 32, 8192, 151, 60
    print(f"case config1: {S} x {P_}, ngrid {G}, ADI, FP64, shot noise, {NST} steps", flush=True)
    steps = api.steps_array(list(fodo(1))[:NST])
    beam = api.load_beam(nslice=S, npart=P_, nbins=4, gamma0=GAMMA0, espread_rel=ESPREAD,
                         emitn=EMITN, beta=BETA, slicelength=LAMBDA0, current=3000.0,
                         shotnoise=True, seed=1234568)
    re = np.zeros(S * G * G)
    im = np.zeros_like(re)
    braw, fraw, b, f = gpu_run(gpu_ctx, api.FP64, api.ADI, beam, G, GM151, steps, re, im,
                               timerun=True, keep=True)
    ob, of, obraw, ofraw = oracle_run(beam, G, GM151, steps, re, im, solver=api.ADI,
                                      timerun=True)
    try:
        # ring state after NST slippage events: bit-exact
        first, accuslip = f.ring()
        assert first == of.first
        assert accuslip == of.s.accuslip
        assert first == (-NST) % S
        # particle counts per slice: exact in every record
        assert np.array_equal(braw[..., 12], obraw[..., 12])
        # particle-to-cell indices of the final beam: the device rule equals
        # the oracle's on the device's positions, and the positions agree
        x, y, px, py, g, th = b.download()
        cell = b.cell_index(f)
        ref = np.empty(x.size, dtype=np.int32)
        _oracle.lib().orc_cell_index(x.size, dp(x), dp(y), GM151, f.dgrid, G, ip(ref))
        assert np.array_equal(cell, ref)
        oref = np.empty(x.size, dtype=np.int32)
        _oracle.lib().orc_cell_index(x.size, dp(ob.x), dp(ob.y), GM151, f.dgrid, G, ip(oref))
        same = np.mean(cell == oref)
        assert relerr(x, ob.x) < 1e-12 and relerr(y, ob.y) < 1e-12
        assert relerr(g - GAMMA0, ob.gamma - GAMMA0) < 1e-10
        P, B = records(braw, fraw, beam.current, G, GM151)
        oP, oB = records(obraw, ofraw, beam.current, G, GM151)
        ep, eb = relerr(P[1:], oP[1:]), relerr(B, oB)
        print(f"  first {first}, cells equal to the oracle's run: {same:.6f}; power {ep:.2e} "
              f"bunching {eb:.2e}", flush=True)
        assert same == 1.0
        assert ep < FP64_TOL and eb < FP64_TOL
    finally:
        f.close()
        b.close()


# ---- configuration 2 ---------------------------------------------------------

def _config2_beam(S, P_):
    """A window of S slices at the centre of the 4096-slice Gaussian bunch
    (peak 4 kA, sigma 1000 slices), with the 1e-4 end-to-end chirp scaled to
    the window."""
    s = np.arange(S, dtype=np.float64) + 2048 - S / 2
    current = 4000.0 * np.exp(-0.5 * ((s - 2047.5) / 1000.0) ** 2)
    return api.load_beam(nslice=S, npart=P_, nbins=4, gamma0=GAMMA0, espread_rel=ESPREAD,
                         emitn=EMITN, beta=BETA, slicelength=LAMBDA0, current=current,
                         shotnoise=True, seed=1234569, chirp_rel=1e-4 * S / 4096)


@pytest.fixture(scope="module")
def config2_oracle():
    S, P_, G, NST = 32, 16384, 256, 40
    steps = api.steps_array(list(fodo(1))[:NST])
    beam = _config2_beam(S, P_)
    re = np.zeros(S * G * G)
    im = np.zeros_like(re)
    t = time.perf_counter()
    ob, of, obraw, ofraw = oracle_run(beam, G, GM256, steps, re, im, solver=api.FFT,
                                      timerun=True)
    print(f"  config2 oracle {time.perf_counter() - t:.1f} s", flush=True)
    return beam, steps, re, im, of, obraw, ofraw


@pytest.mark.parametrize("prec", [api.FP64, api.MIXED])
def test_config2_fft_run(gpu_ctx, config2_oracle, prec):
    beam, steps, re, im, of, obraw, ofraw = config2_oracle
    S, G = beam.offsets.size - 1, 256
    print(f"case config2: {S} x 16384, ngrid {G}, FFT, prec {prec}, {len(steps)} steps",
          flush=True)
    braw, fraw, b, f = gpu_run(gpu_ctx, prec, api.FFT, beam, G, GM256, steps, re, im,
                               timerun=True, keep=True)
    try:
        assert f.ring()[0] == of.first
        assert np.array_equal(braw[..., 12], obraw[..., 12])
        P, B = records(braw, fraw, beam.current, G, GM256)
        oP, oB = records(obraw, ofraw, beam.current, G, GM256)
        bq = [api.finalize_beam(np.ravel(braw[r]), 1, beam.current, GAMMA0)[0][:, 1]
              for r in range(braw.shape[0])]
        obq = [api.finalize_beam(np.ravel(obraw[r]), 1, beam.current, GAMMA0)[0][:, 1]
               for r in range(obraw.shape[0])]
        ep, eb, es = relerr(P[1:], oP[1:]), relerr(B, oB), relerr(bq, obq)
        print(f"  power {ep:.2e} bunching {eb:.2e} energy spread {es:.2e}", flush=True)
        tol = FP64_TOL if prec == api.FP64 else MIXED_TOL
        assert ep < tol and eb < tol and es < tol
    finally:
        f.close()
        b.close()


# ---- configuration 3 ---------------------------------------------------------

def _config3_runs(nmod):
    """8 tapered steady-state runs covering every grid of the scan twice and
    every particle count, aw(z) = aw0 (1 - a z - b z^2) per module with
    (a, b) drawn from [0, 2e-3] (seed 1234570)."""
    rng = np.random.default_rng(1234570)
    grids = [101, 151, 201, 255, 255, 201, 151, 101]
    parts = [4096, 8192, 16384, 4096, 8192, 16384, 8192, 4096]
    z = np.floor(np.arange(100 * nmod) / 100) / nmod
    runs = []
    for g, p in zip(grids, parts):
        a, b = rng.uniform(0.0, 2e-3, 2)
        runs.append(dict(npart=p, nbins=4, ngrid=g, gridmax=GM151, current=3000.0,
                         gamma0=GAMMA0, espread_rel=ESPREAD, emitn=EMITN, beta=BETA,
                         shotnoise=False, seed=int(rng.integers(1, 1 << 31)), seed_power=1e3,
                         seed_w0=50e-6, aw=AW * (1 - a * z - b * z * z)))
    return runs


def _config3_oracle_one(r, steps):
    # one oracle thread per run (OpenMP's thread count is per calling thread)
    _oracle.lib().orc_set_threads(1)
    beam = api.load_beam(nslice=1, npart=r["npart"], nbins=4, gamma0=GAMMA0,
                         espread_rel=ESPREAD, emitn=EMITN, beta=BETA, slicelength=LAMBDA0,
                         current=r["current"], shotnoise=False, seed=r["seed"])
    st = api.steps_array(list(steps))
    for i in range(len(st)):
        st[i].aw = float(r["aw"][i])
    re, im = api.seed_gaussian(1, r["ngrid"], GM151, LAMBDA0, r["seed_power"], r["seed_w0"])
    _, _, obraw, ofraw = oracle_run(beam, r["ngrid"], GM151, st, re, im, solver=api.ADI,
                                    timerun=False)
    return records(obraw, ofraw, beam.current, r["ngrid"], GM151)


@pytest.fixture(scope="module")
def config3_oracle():
    nmod = 10
    steps = fodo(nmod, output_step=10)
    runs = _config3_runs(nmod)
    t = time.perf_counter()
    with cf.ThreadPoolExecutor(max_workers=len(runs)) as ex:
        ref = list(ex.map(lambda r: _config3_oracle_one(r, steps), runs))
    print(f"  config3 oracle {time.perf_counter() - t:.1f} s", flush=True)
    return runs, steps, ref


@pytest.mark.parametrize("prec", [api.FP64, api.MIXED])
def test_config3_taper_batch(gpu_ctx, config3_oracle, prec):
    runs, steps, ref = config3_oracle
    print(f"case config3: batch of {len(runs)} tapered runs, ngrid 101-255, "
          f"{len(steps)} steps, prec {prec}", flush=True)
    batch = api.Batch(gpu_ctx, prec, api.ADI, runs, steps, LAMBDA0)
    try:
        assert batch.advance(len(steps)) == len(steps)
        tol = FP64_TOL if prec == api.FP64 else MIXED_TOL
        for i, (r, (oP, oB)) in enumerate(zip(runs, ref)):
            braw, fraw = batch.download(i)
            P, B = records(braw[:, None, :], fraw[:, None, :], [r["current"]], r["ngrid"], GM151)
            bfloor = 1e-6 if prec == api.FP64 else shot_noise_bunching(r["current"])
            ep, eb = per_record(P, oP), per_record(B, oB, floor=bfloor)
            print(f"  run {i} (ngrid {r['ngrid']}, {r['npart']} particles): power gain "
                  f"{oP[-1, 0] / oP[0, 0]:.2e}, worst record power {ep:.2e} bunching {eb:.2e}",
                  flush=True)
            assert ep < tol and eb < tol, i
    finally:
        batch.close()


# ---- configuration 4 ---------------------------------------------------------

def _config4_stream(n, seed):
    """n requests: 70 % steady state (grid 101/151/201/255, 4096/8192/16384
    particles, linear + quadratic taper from [0, 2e-3]) and 30 %
    time-dependent (128-512 slices, grid 101-255, 4096-16384
    particles/slice, shot noise), on one 100-period module, ADI, with
    Poisson inter-arrival times of mean 20 ms."""
    rng = np.random.default_rng(seed)
    steps = api.lattice_fodo(LAMBDAU, [AW], 100, 46, 3, 1.5, 10)
    z = np.arange(1, len(steps) + 1) / len(steps)
    out = []
    for _ in range(n):
        gap = rng.exponential(0.02)
        if rng.random() < 0.7:
            a, b = rng.uniform(0.0, 2e-3, 2)
            r = dict(kind=0, nslice=1, ngrid=int(rng.choice([101, 151, 201, 255])),
                     npart=int(rng.choice([4096, 8192, 16384])), aw=AW * (1 - a * z - b * z * z),
                     seed=int(rng.integers(1, 1 << 31)), seed_power=1e3)
        else:
            r = dict(kind=1, nslice=int(rng.integers(128, 513)),
                     ngrid=int(rng.integers(101, 256)), npart=4 * int(rng.integers(1024, 4097)),
                     seed=int(rng.integers(1, 1 << 31)), seed_power=0.0)
        out.append((gap, r))
    return steps, out


def _alone(ctx, prec, r, steps):
    """The request run alone through g4h_run from the state the server
    builds for it: steady-state beams from the host loader, time-dependent
    beams generated on the device with the seed plane in every slice."""
    S, G = r["nslice"], r["ngrid"]
    kw = dict(nslice=S, npart=r["npart"], nbins=4, gamma0=GAMMA0, espread_rel=ESPREAD,
              emitn=EMITN, beta=BETA, slicelength=LAMBDA0, current=3000.0, shotnoise=True,
              seed=r["seed"])
    st = api.steps_array(list(steps))
    if r.get("aw") is not None:
        for i in range(len(st)):
            st[i].aw = float(r["aw"][i])
    if r["seed_power"] > 0:
        re, im = api.seed_gaussian(1, G, GM151, LAMBDA0, r["seed_power"], 50e-6)
    else:
        re = np.zeros(G * G)
        im = np.zeros_like(re)
    off = (np.arange(S + 1) * r["npart"]).astype(np.int32)
    b = api.Beam(ctx, prec, off, np.full(S, 3000.0), GAMMA0)
    f = api.Field(ctx, prec, S, G, GM151, LAMBDA0, LAMBDA0)
    if r["kind"] == 1:
        b.generate(**kw)
        f.fill_planes(re, im, 0)
        beam = None
    else:
        beam = api.load_beam(**kw)
        b.upload(beam.x, beam.y, beam.px, beam.py, beam.gamma, beam.theta)
        f.upload(re, im, 0)
    f.init_solver(api.ADI)
    run = api.Run(ctx, b, f, st, LAMBDA0, 1.0, r["kind"] == 1, False, 1)
    run.advance(len(st))
    out = run.download_diag()
    run.close()
    if r["kind"] == 1:
        # the generated beam, for the oracle
        beam = api.load_beam(**kw)
    f.close()
    b.close()
    return out, beam, st, re, im


@pytest.mark.parametrize("prec", [api.FP64, api.MIXED])
def test_config4_serving_stream(gpu_ctx, prec):
    steps, stream = _config4_stream(16, 1234571)
    kinds = [r["kind"] for _, r in stream]
    assert 0 in kinds and 1 in kinds
    print(f"case config4: {len(stream)} requests ({kinds.count(1)} time-dependent), "
          f"prec {prec}", flush=True)
    srv = api.Server(0, hbm_budget=200 * 10 ** 9, nstreams=4)
    ids = []
    try:
        for gap, r in stream:
            time.sleep(gap)
            ids.append(srv.submit(precision=prec, solver=api.ADI, steps=steps, gridmax=GM151,
                                  current=3000.0, gamma0=GAMMA0, lambda0=LAMBDA0,
                                  espread_rel=ESPREAD, **r))
        results = [srv.result(i) for i in ids]
        st = srv.stats()
        assert st["completed"] == len(stream)
        assert 0 < st["p50_s"] <= st["p99_s"]
        for i in ids:
            srv.release(i)
    finally:
        srv.close()
    tol = 1e-12 if prec == api.FP64 else 1e-6
    # the smallest request of each kind also goes to the oracle
    check = {}
    for k in (0, 1):
        cand = [j for j, (_, r) in enumerate(stream) if r["kind"] == k]
        check[k] = min(cand, key=lambda j: stream[j][1]["nslice"] * stream[j][1]["npart"]
                       * stream[j][1]["ngrid"])
    for j, ((_, r), (braw, fraw)) in enumerate(zip(stream, results)):
        (lb, lf), beam, st_, re, im = _alone(gpu_ctx, prec, r, steps)
        assert braw.shape == lb.shape and fraw.shape == lf.shape
        P, B = records(braw, fraw, np.full(r["nslice"], 3000.0), r["ngrid"], GM151)
        lP, lB = records(lb, lf, np.full(r["nslice"], 3000.0), r["ngrid"], GM151)
        ep, eb = relerr(P, lP), relerr(B, lB)
        msg = f"  request {j} kind {r['kind']} {r['nslice']}x{r['npart']} ngrid {r['ngrid']}: " \
              f"in mix vs alone power {ep:.2e} bunching {eb:.2e}"
        if j == check[r["kind"]]:
            S = r["nslice"]
            if r["kind"] == 1:
                re = np.tile(re, S)
                im = np.tile(im, S)
            _, _, obraw, ofraw = oracle_run(beam, r["ngrid"], GM151, st_, re, im,
                                            solver=api.ADI, timerun=r["kind"] == 1)
            oP, oB = records(obraw, ofraw, np.full(S, 3000.0), r["ngrid"], GM151)
            op_, ob_ = relerr(P[1:], oP[1:]), relerr(B, oB)
            msg += f"; vs oracle power {op_:.2e} bunching {ob_:.2e}"
            otol = FP64_TOL if prec == api.FP64 else MIXED_TOL
            print(msg, flush=True)
            assert op_ < otol and ob_ < otol, j
        else:
            print(msg, flush=True)
        assert ep < tol and eb < tol, j
