"""GPU: the request loop (g4h_server_*).

A mix of steady-state requests (batched together) and time-dependent
requests is submitted at once to a server with three stream lanes, in each
precision; every result must equal the same request run alone through
g4h_run (FP64 1e-12, MIXED 1e-6 relative on slice power and bunching: the
batched MIXED kernels sum in a different order), the latency statistics must
cover every request, and a request larger than the HBM budget must fail
with G4H_ERR_OUT_OF_MEMORY without disturbing the others."""
from __future__ import annotations

import numpy as np
import pytest

from cases import AW, GAMMA0, LAMBDA0, LAMBDAU
from genesis4_cuda_port_amd import api
from genesis4_cuda_port_amd._abi import G4Error

pytestmark = pytest.mark.gpu

GRIDMAX = 0.3e-3


def relerr(a, b):
    den = np.linalg.norm(np.ravel(b))
    return np.linalg.norm(np.ravel(a) - np.ravel(b)) / (den if den > 0 else 1.0)


def _alone(ctx, r, steps, prec):
    """The request run alone through g4h_run, from the state the server
    builds: steady-state beams from the host loader (as g4h_batch_create),
    time-dependent ones generated on the device with the seed plane filled
    into every slice (as the server does for them)."""
    kw = dict(nslice=r["nslice"], npart=r["npart"], nbins=4, gamma0=GAMMA0, espread_rel=1e-4,
              emitn=0.4e-6, beta=15.0, slicelength=LAMBDA0 * r.get("sample", 1.0),
              current=r["current"], shotnoise=True, seed=r["seed"])
    st = api.steps_array(list(steps))
    if r.get("aw") is not None:
        for i in range(len(st)):
            st[i].aw = float(r["aw"][i])
    S, G = r["nslice"], r["ngrid"]
    if r.get("seed_power", 0) > 0:
        re, im = api.seed_gaussian(1 if r["kind"] == 1 else S, G, GRIDMAX, LAMBDA0,
                                   r["seed_power"], 50e-6)
    else:
        re = np.zeros((1 if r["kind"] == 1 else S) * G * G)
        im = np.zeros_like(re)
    off = (np.arange(S + 1) * r["npart"]).astype(np.int32)
    b = api.Beam(ctx, prec, off, np.full(S, r["current"]), GAMMA0)
    f = api.Field(ctx, prec, S, G, GRIDMAX, LAMBDA0, LAMBDA0 * r.get("sample", 1.0))
    if r["kind"] == 1:
        b.generate(**kw)
        f.fill_planes(re, im, 0)
    else:
        beam = api.load_beam(**kw)
        b.upload(beam.x, beam.y, beam.px, beam.py, beam.gamma, beam.theta)
        f.upload(re, im, 0)
    f.init_solver(r["solver"])
    run = api.Run(ctx, b, f, st, LAMBDA0, r.get("sample", 1.0), r["kind"] == 1, False, 1)
    run.advance(len(st))
    out = run.download_diag()
    run.close()
    f.close()
    b.close()
    return out


@pytest.mark.parametrize("prec", [api.FP64, api.MIXED])
def test_server_mix_matches_lone_runs(gpu_ctx, prec):
    print(f"case server mix prec={prec}", flush=True)
    steps = api.lattice_fodo(LAMBDAU, [AW], 24, 10, 3, 1.5, 2)
    taper = AW * (1 - 1e-3 * np.arange(len(steps)) / len(steps))
    common = dict(precision=prec, gridmax=GRIDMAX, gamma0=GAMMA0, lambda0=LAMBDA0,
                  espread_rel=1e-4)
    reqs = [
        dict(kind=0, solver=api.ADI, nslice=1, npart=512, ngrid=33, current=3000.0, seed=21,
             seed_power=1e6),
        dict(kind=0, solver=api.ADI, nslice=1, npart=1024, ngrid=31, current=2000.0, seed=22),
        dict(kind=1, solver=api.FFT, nslice=9, npart=256, ngrid=32, current=3000.0, seed=23),
        dict(kind=0, solver=api.ADI, nslice=1, npart=256, ngrid=33, current=3000.0, seed=24,
             aw=taper),
        dict(kind=1, solver=api.ADI, nslice=6, npart=512, ngrid=31, current=2500.0, seed=25,
             seed_power=1e5),
        dict(kind=0, solver=api.ADI, nslice=1, npart=768, ngrid=31, current=3000.0, seed=26,
             seed_power=1e5, aw=taper[::-1].copy()),
    ]
    srv = api.Server(0, hbm_budget=8 << 30, nstreams=3)
    ids = [srv.submit(steps=steps, **common, **r) for r in reqs]
    # a request that can never fit the budget fails on its own
    big = srv.submit(steps=steps, kind=1, solver=api.FFT, nslice=256, npart=1 << 20, ngrid=512,
                     current=3000.0, seed=1, **common)
    results = [srv.result(i) for i in ids]
    with pytest.raises(G4Error) as e:
        srv.wait(big)
    assert e.value.status == 3  # G4H_ERR_OUT_OF_MEMORY
    st = srv.stats()
    assert st["completed"] == len(reqs)
    assert 0 < st["p50_s"] <= st["p99_s"]
    for i in ids:
        srv.release(i)
    srv.close()
    for k, (r, (braw, fraw)) in enumerate(zip(reqs, results)):
        lb, lf = _alone(gpu_ctx, r, steps, prec)
        assert braw.shape == lb.shape and fraw.shape == lf.shape
        eb, ef = relerr(braw, lb), relerr(fraw, lf)
        print(f"  request {k}: beam {eb:.3e} field {ef:.3e}", flush=True)
        tol = 1e-12 if prec == api.FP64 else 1e-6
        assert eb < tol and ef < tol, k
