"""GPU: g4h_diag_beam_raw_noalloc, the beam-sums entry point that allocates no
device memory (its scratch is one fixed array per GPU).

Bar: bit-for-bit equal to g4h_diag_beam_raw on the same beam, on two calls
in a row (the second with a smaller beam and fewer harmonics), in FP64 and
MIXED, for a beam whose pass is split into several workgroups per slice
and one that is not; a beam beyond the array's size is refused with
G4H_ERR_INVALID_ARGUMENT before any work."""
from __future__ import annotations

import numpy as np
import pytest

import cases
from cases import GAMMA0
from genesis4_cuda_port_amd import api
from genesis4_cuda_port_amd._abi import G4Error

pytestmark = pytest.mark.gpu


def _beam(ctx, prec, nslice, npart, seed):
    c = cases.make_case(nslice=nslice, npart=npart, ngrid=33, seed=seed)
    b = api.Beam(ctx, prec, c.beam.offsets, c.beam.current, GAMMA0)
    b.upload(c.beam.x, c.beam.y, c.beam.px, c.beam.py, c.beam.gamma, c.beam.theta)
    return b


@pytest.mark.parametrize("prec", [api.FP64, api.MIXED])
def test_noalloc_beam_sums_equal_allocating_entry_point(gpu_ctx, prec):
    # 24 x 16384: split pass (several workgroups per slice, partials and
    # counters in the array); then 640 x 1024: one workgroup per slice
    for nslice, npart, nharm in ((24, 16384, 4), (640, 1024, 2), (5, 8192, 1)):
        print(f"case noalloc prec={prec} {nslice}x{npart} nharm={nharm}", flush=True)
        b = _beam(gpu_ctx, prec, nslice, npart, seed=100 + nslice)
        try:
            want = b.diag_raw(nharm)
            got = b.diag_raw_noalloc(nharm)
            assert got.shape == want.shape
            assert np.array_equal(got, want), np.max(np.abs(got - want))
            # a second call on the same beam gives the same sums again
            assert np.array_equal(b.diag_raw_noalloc(nharm), want)
        finally:
            b.close()


def test_noalloc_refuses_beams_beyond_its_array(gpu_ctx):
    print("case noalloc refuses oversize", flush=True)
    nslice = 60000  # 60000 x 39 doubles > 8 MiB at nharm = 8
    off = np.arange(nslice + 1, dtype=np.int32)  # one particle per slice
    b = api.Beam(gpu_ctx, api.FP64, off, np.full(nslice, 3000.0), GAMMA0)
    z = np.zeros(nslice)
    b.upload(z, z, z, z, np.full(nslice, GAMMA0), np.linspace(0.0, 6.0, nslice))
    try:
        with pytest.raises(G4Error) as e:
            b.diag_raw_noalloc(8)
        assert e.value.status == 1  # G4H_ERR_INVALID_ARGUMENT
        assert "scratch" in str(e.value)
    finally:
        b.close()
    # 30000 slices at one harmonic fit (30000 x 25 doubles < 8 MiB)
    nslice = 30000
    off = np.arange(nslice + 1, dtype=np.int32)
    b = api.Beam(gpu_ctx, api.FP64, off, np.full(nslice, 3000.0), GAMMA0)
    z = np.zeros(nslice)
    b.upload(z, z, z, z, np.full(nslice, GAMMA0), np.linspace(0.0, 6.0, nslice))
    try:
        assert np.array_equal(b.diag_raw_noalloc(1), b.diag_raw(1))
    finally:
        b.close()
