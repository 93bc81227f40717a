"""GPU: checkpoint and resume of a time-dependent run.  The state of a run
is the beam arrays, the field planes and the field's ring state (first,
accuslip); downloading them after k steps and continuing from an upload in
a fresh context with the remaining lattice steps must reproduce the
uninterrupted run (records and final state), as a restart from the
reference's dumps would."""
from __future__ import annotations

import numpy as np
import pytest

import cases
from cases import LAMBDA0, GAMMA0
from genesis4_cuda_port_amd import api

pytestmark = pytest.mark.gpu

from test_gpu_parity import relerr  # noqa: E402


def _fresh(ctx, c, prec, kind, state=None, field=None, ring=None):
    b = api.Beam(ctx, prec, c.beam.offsets, c.beam.current, GAMMA0)
    b.upload(*(state if state is not None else (c.beam.x, c.beam.y, c.beam.px, c.beam.py,
                                                 c.beam.gamma, c.beam.theta)))
    f = api.Field(ctx, prec, c.nslice, c.ngrid, c.gridmax, LAMBDA0, c.slicelength)
    if field is None:
        f.upload(c.seed_re, c.seed_im, 0)
    else:
        f.upload(field[0], field[1], field[2])
    if ring is not None:
        f.set_ring(*ring)
    f.init_solver(kind)
    return b, f


@pytest.mark.parametrize("prec", [api.FP64, api.MIXED])
@pytest.mark.parametrize("kind", [api.ADI, api.FFT])
def test_resume_matches_uninterrupted(prec, kind):
    print(f"case checkpoint prec={prec} kind={kind}", flush=True)
    c = cases.make_case(nslice=10, npart=1024, ngrid=32 if kind == api.FFT else 33, nsteps=14,
                        seed_power=1e5)
    k = 6
    with api.Context(0) as ctx:
        b, f = _fresh(ctx, c, prec, kind)
        run = api.Run(ctx, b, f, c.steps, LAMBDA0, c.sample, c.timerun, False, 1)
        run.advance(len(c.steps))
        ref_b, ref_f = run.download_diag()
        ref_state = b.download()
        run.close()
        f.close()
        b.close()
    with api.Context(0) as ctx:
        b, f = _fresh(ctx, c, prec, kind)
        run = api.Run(ctx, b, f, c.steps[:k], LAMBDA0, c.sample, c.timerun, False, 1)
        run.advance(k)
        state = b.download()
        field = f.download()
        ring = f.ring()
        run.close()
        f.close()
        b.close()
    assert ring[0] == field[2]
    with api.Context(0) as ctx:
        b, f = _fresh(ctx, c, prec, kind, state=state, field=field, ring=ring)
        run = api.Run(ctx, b, f, c.steps[k:], LAMBDA0, c.sample, c.timerun, False, 1)
        run.advance(len(c.steps) - k)
        got_b, got_f = run.download_diag()
        got_state = b.download()
        end_ring = f.ring()
        run.close()
        f.close()
        b.close()
    tol = 1e-12 if prec == api.FP64 else 1e-5
    # the resumed run's first record is the state after k steps
    assert got_b.shape[0] == ref_b.shape[0] - k
    eb, ef = relerr(got_b, ref_b[k:]), relerr(got_f, ref_f[k:])
    print(f"  beam raw {eb:.2e} field raw {ef:.2e} ring {end_ring}", flush=True)
    assert eb < tol and ef < tol
    assert np.array_equal(got_b[:, :, 12], ref_b[k:, :, 12])
    for x, y in zip(got_state, ref_state):
        assert relerr(x, y) < tol
