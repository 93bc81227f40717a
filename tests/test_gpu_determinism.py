"""GPU: run-to-run reproducibility.  The deposition adds into LDS with
atomics, whose order is not fixed, so two runs of the same input may differ
at round-off in the sums but must agree far inside the parity tolerances;
particle counts and ring offsets are integers and must agree exactly."""
from __future__ import annotations

import numpy as np
import pytest

import cases
from cases import LAMBDA0
from genesis4_cuda_port_amd import api

pytestmark = pytest.mark.gpu

from test_gpu_parity import gpu_beam, gpu_field, relerr  # noqa: E402


def _run(prec, kind, c):
    with api.Context(0) as ctx:
        b = gpu_beam(ctx, c, prec)
        f = gpu_field(ctx, c, prec, kind=kind)
        run = api.Run(ctx, b, f, c.steps, LAMBDA0, c.sample, c.timerun, False, 2)
        run.advance(len(c.steps))
        braw, fraw = run.download_diag()
        state = b.download()
        ring = f.download()[2]
        run.close()
        f.close()
        b.close()
    return braw, fraw, state, ring


@pytest.mark.parametrize("prec", [api.FP64, api.MIXED])
@pytest.mark.parametrize("kind", [api.ADI, api.FFT])
def test_two_runs_agree(prec, kind):
    print(f"case determinism prec={prec} kind={kind}", flush=True)
    c = cases.make_case(nslice=12, npart=2048, ngrid=32 if kind == api.FFT else 33, nsteps=16,
                        seed_power=1e4)
    a = _run(prec, kind, c)
    b = _run(prec, kind, c)
    tol = 1e-13 if prec == api.FP64 else 1e-6
    eb, ef = relerr(a[0], b[0]), relerr(a[1], b[1])
    print(f"  beam raw {eb:.2e} field raw {ef:.2e}", flush=True)
    assert eb < tol and ef < tol
    assert np.array_equal(a[0][:, :, 12], b[0][:, :, 12])  # particle counts
    assert a[3] == b[3]  # ring offset
    for x, y in zip(a[2], b[2]):
        assert relerr(x, y) < tol
