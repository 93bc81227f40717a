"""GPU: the quiet-start beam generated on the device (g4h_beam_generate)
and the seed field filled plane by plane (g4h_field_fill_planes).

The device generator runs the host loader's routine (loader.hpp), so its
particles equal g4h_load_beam's up to the last bits of the FP64 library
functions (inverse normal CDF, Box-Muller): coordinates within 1e-14
relative, the nbins mirrored particles of a beamlet share their transverse
coordinates and energy exactly.  MIXED stores FP32 coordinates and
gamma - gamma_ref, the host values rounded."""
from __future__ import annotations

import numpy as np
import pytest

from cases import GAMMA0, LAMBDA0
from genesis4_cuda_port_amd import api

pytestmark = pytest.mark.gpu

KW = dict(nslice=9, npart=2048, nbins=4, gamma0=GAMMA0, espread_rel=1e-4, emitn=0.4e-6, beta=15.0,
          alpha=0.3, slicelength=LAMBDA0, shotnoise=True, seed=77, chirp_rel=2e-4,
          emitn_y=0.5e-6, beta_y=12.0, alpha_y=-0.2)


def _relmax(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


@pytest.mark.parametrize("prec", [api.FP64, api.MIXED])
def test_device_generator_matches_host_loader(gpu_ctx, prec):
    print(f"case device_generator prec={prec}", flush=True)
    cur = np.linspace(1000.0, 4000.0, KW["nslice"])
    host = api.load_beam(current=cur, **KW)
    b = api.Beam(gpu_ctx, prec, host.offsets, cur, GAMMA0)
    b.generate(current=cur, **KW)
    x, y, px, py, g, th = b.download()
    if prec == api.FP64:
        for got, ref in ((x, host.x), (y, host.y), (px, host.px), (py, host.py)):
            assert _relmax(got, ref) < 1e-14
        assert _relmax(g - GAMMA0, host.gamma - GAMMA0) < 1e-12
        assert np.max(np.abs(th - host.theta)) < 1e-13
    else:
        for got, ref in ((x, host.x), (y, host.y), (px, host.px), (py, host.py)):
            assert _relmax(got, ref) < 1e-6
        assert _relmax(g - GAMMA0, host.gamma - GAMMA0) < 1e-5
        assert np.max(np.abs(th - host.theta)) < 1e-13
    for a in (x, px, y, py, g):
        v = a.reshape(-1, 4)
        assert np.all(v == v[:, :1])
    b.close()


def test_fill_planes_equals_upload(gpu_ctx):
    print("case fill_planes", flush=True)
    S, G = 5, 33
    re, im = api.seed_gaussian(S, G, 0.3e-3, LAMBDA0, 1e4, 40e-6)
    a = api.Field(gpu_ctx, api.FP64, S, G, 0.3e-3, LAMBDA0, LAMBDA0)
    a.upload(re, im, 2)
    b = api.Field(gpu_ctx, api.FP64, S, G, 0.3e-3, LAMBDA0, LAMBDA0)
    b.fill_planes(re[:G * G], im[:G * G], 2)
    ra, ia, fa = a.download()
    rb, ib, fb = b.download()
    assert fa == fb == 2 and np.array_equal(ra, rb) and np.array_equal(ia, ib)
    a.close()
    b.close()
