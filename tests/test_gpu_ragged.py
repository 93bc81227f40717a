"""GPU: time-dependent runs over ragged beams — uneven particle counts per
slice, empty slices, and slices with particles but no current — through the
fused run loop against the FP64 oracle, with one and with three workgroups
per slice.

Bars: particle counts of every slice bit-exact; FP64 slice power and
bunching (slices that hold particles) within 1e-10 relative; the empty
slices' beam sums exactly zero."""
from __future__ import annotations

import dataclasses

import numpy as np
import pytest

import _oracle
from _oracle import OField
import cases
from cases import GAMMA0, LAMBDA0
from genesis4_cuda_port_amd import api
from genesis4_cuda_port_amd._abi import TUNE_SPLIT

pytestmark = pytest.mark.gpu

from test_gpu_parity import FP64_TOL, _power_bunching, gpu_beam, gpu_field, relerr  # noqa: E402

COUNTS = [512, 0, 256, 1024, 4, 0, 768, 1020]
CURRENT = [3000.0, 0.0, 2000.0, 3500.0, 3000.0, 1000.0, 0.0, 2500.0]


def ragged_case(kind):
    c = cases.make_case(nslice=len(COUNTS), npart=1024, ngrid=32 if kind == api.FFT else 31,
                        nsteps=16, seed_power=1e4, current=CURRENT)
    b = c.beam
    sel = np.concatenate([np.arange(s * 1024, s * 1024 + n) for s, n in enumerate(COUNTS)])
    off = np.concatenate([[0], np.cumsum(COUNTS)]).astype(np.int32)
    beam = api.BeamArrays(b.x[sel].copy(), b.y[sel].copy(), b.px[sel].copy(), b.py[sel].copy(),
                          b.gamma[sel].copy(), b.theta[sel].copy(), off, b.current.copy())
    return dataclasses.replace(c, beam=beam)


@pytest.mark.parametrize("kind", [api.ADI, api.FFT])
@pytest.mark.parametrize("split", [1, 3])
def test_ragged_run_fp64(kind, split):
    print(f"case ragged kind={kind} split={split}", flush=True)
    c = ragged_case(kind)
    with api.Context(0) as ctx:
        ctx.set_variant(TUNE_SPLIT, split)
        b = gpu_beam(ctx, c)
        f = gpu_field(ctx, c, kind=kind)
        run = api.Run(ctx, b, f, c.steps, LAMBDA0, c.sample, c.timerun, False, 2)
        run.advance(len(c.steps))
        braw, fraw = run.download_diag()
        run.close()
        f.close()
        b.close()
    of = OField(c.nslice, c.ngrid, c.gridmax, LAMBDA0, c.slicelength, 1, c.seed_re, c.seed_im)
    ob, obraw, ofraw = _oracle.run(c.beam, of, c.steps, gamma0=GAMMA0, lambda0=LAMBDA0,
                                   solver=kind, sample=c.sample, timerun=c.timerun, nharm=2)
    assert braw.shape == obraw.shape
    cnt = braw[:, :, 12]
    assert np.array_equal(cnt, obraw[:, :, 12])
    assert np.array_equal(cnt[0], np.array(COUNTS, dtype=float))
    empty = np.array(COUNTS) == 0
    assert np.all(braw[:, empty, :12] == 0.0)

    class _F:
        dgrid = 2 * c.gridmax / (c.ngrid - 1)
        xlambda = LAMBDA0
        slicelength = c.slicelength

    P, B = _power_bunching(braw, fraw, c, _F)
    oP, oB = _power_bunching(obraw, ofraw, c, _F)
    ep = relerr(P, oP)
    eb = relerr(B[:, ~empty], oB[:, ~empty])
    print(f"  power {ep:.3e} bunching {eb:.3e}", flush=True)
    assert ep < FP64_TOL and eb < FP64_TOL
