"""GPU: ensembles replayed from a captured hipGraph (g4h_run_capture /
g4h_run_replay).

One run is advanced eagerly and captured; it is then replayed several
times, alternating between two ensemble members (beam and seed field of
different seeds), with eager work and synchronisations in between.  Every
replay must give the records of the same member run eagerly in a fresh
context: FP64 within 1e-12 relative (the deposition sums may add in another
order), MIXED within 1e-5.  The first replay is also checked against the
FP64 oracle run of its member (slice power and bunching: FP64 1e-10, MIXED
1e-3).  Replaying without re-initialising the field is refused."""
from __future__ import annotations

import numpy as np
import pytest

import _oracle
from _oracle import OField
import cases
from cases import GAMMA0, LAMBDA0, quad_step
from genesis4_cuda_port_amd import api
from genesis4_cuda_port_amd._abi import G4Error

pytestmark = pytest.mark.gpu

from test_gpu_parity import (FP64_TOL, MIXED_TOL, _power_bunching, gpu_beam,  # noqa: E402
                             gpu_field, relerr)


def _eager(prec, kind, c):
    with api.Context(0) as ctx:
        b = gpu_beam(ctx, c, prec)
        f = gpu_field(ctx, c, prec, kind=kind)
        run = api.Run(ctx, b, f, c.steps, LAMBDA0, c.sample, c.timerun, False, 2)
        run.advance(len(c.steps))
        out = run.download_diag()
        run.close()
        f.close()
        b.close()
    return out


@pytest.mark.parametrize("prec,kind", [(api.FP64, api.ADI), (api.FP64, api.FFT),
                                       (api.MIXED, api.FFT), (api.MIXED, api.ADI)])
def test_graph_replay_matches_eager(prec, kind):
    print(f"case graph_replay prec={prec} kind={kind}", flush=True)
    g = 32 if kind == api.FFT else 31
    members = [cases.make_case(nslice=8, npart=1024, ngrid=g, nsteps=20, seed_power=1e4,
                               seed=seed) for seed in (1234568, 4321)]
    members[1].seed_re *= 0.5
    ref = [_eager(prec, kind, m) for m in members]
    tol = 1e-12 if prec == api.FP64 else 1e-5
    with api.Context(0) as ctx:
        c = members[0]
        b = gpu_beam(ctx, c, prec)
        f = gpu_field(ctx, c, prec, kind=kind)
        run = api.Run(ctx, b, f, c.steps, LAMBDA0, c.sample, c.timerun, False, 2)
        run.advance(len(c.steps))
        first = run.download_diag()
        assert relerr(first[0], ref[0][0]) < tol and relerr(first[1], ref[0][1]) < tol
        run.capture()
        with pytest.raises(G4Error):  # the field still holds the end of the run
            run.replay()
        for rep, k in enumerate((1, 0, 1)):
            m = members[k]
            b.upload(m.beam.x, m.beam.y, m.beam.px, m.beam.py, m.beam.gamma, m.beam.theta)
            f.upload(m.seed_re, m.seed_im, 0)
            run.replay()
            braw, fraw = run.download_diag()
            eb, ef = relerr(braw, ref[k][0]), relerr(fraw, ref[k][1])
            print(f"  replay {rep} (member {k}): beam {eb:.3e} field {ef:.3e}", flush=True)
            assert eb < tol and ef < tol
            if rep == 0:
                of = OField(m.nslice, m.ngrid, m.gridmax, LAMBDA0, m.slicelength, 1, m.seed_re,
                            m.seed_im)
                _, obraw, ofraw = _oracle.run(m.beam, of, m.steps, gamma0=GAMMA0,
                                              lambda0=LAMBDA0, solver=kind, sample=m.sample,
                                              timerun=m.timerun, nharm=2)
                P, B = _power_bunching(braw, fraw, m, f)
                oP, oB = _power_bunching(obraw, ofraw, m, f)
                ep, eb2 = relerr(P[1:], oP[1:]), relerr(B, oB)
                print(f"  replay vs oracle: power {ep:.3e} bunching {eb2:.3e}", flush=True)
                otol = FP64_TOL if prec == api.FP64 else MIXED_TOL
                assert ep < otol and eb2 < otol
            # eager work on the same stream between replays
            b.diag_raw(2)
            ctx.synchronize()
        run.close()
        f.close()
        b.close()


@pytest.mark.parametrize("kind", [api.ADI, api.FFT])
def test_replay_refused_after_table_eviction(kind):
    """A field keeps solver tables for at most 64 distinct step lengths.  When
    another run on the same field steps through 65 of them the tables are
    released and rebuilt, so a graph captured earlier (whose nodes point at the
    released tables) must refuse to replay; capturing a run with 65 distinct
    step lengths fails with a status and leaves the stream usable."""
    print(f"case replay_after_eviction kind={kind}", flush=True)
    g = 32 if kind == api.FFT else 31
    c = cases.make_case(nslice=4, npart=256, ngrid=g, nsteps=6, seed_power=1e4)
    many = api.steps_array([quad_step(delz=cases.LAMBDAU * (1.0 + 1e-3 * k)) for k in range(65)])
    with api.Context(0) as ctx:
        b = gpu_beam(ctx, c, api.FP64)
        f = gpu_field(ctx, c, api.FP64, kind=kind)
        run = api.Run(ctx, b, f, c.steps, LAMBDA0, c.sample, c.timerun, False, 1)
        run.advance(len(c.steps))
        run.capture()
        f.upload(c.seed_re, c.seed_im, 0)
        run.replay()  # tables still live: allowed
        other = api.Run(ctx, b, f, many, LAMBDA0, c.sample, c.timerun, False, 1)
        assert other.advance(65) == 65
        f.upload(c.seed_re, c.seed_im, 0)
        with pytest.raises(G4Error):
            run.replay()
        other.close()
        f.upload(c.seed_re, c.seed_im, 0)
        third = api.Run(ctx, b, f, many, LAMBDA0, c.sample, c.timerun, False, 1)
        assert third.advance(65) == 65
        with pytest.raises(G4Error):
            third.capture()
        third.close()
        # the stream is out of capture mode: eager work still runs
        b.diag_raw(1)
        ctx.synchronize()
        run.close()
        f.close()
        b.close()
