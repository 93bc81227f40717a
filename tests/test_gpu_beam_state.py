"""GPU: beam and field state entry points: beam info, per-slice current
(which scales the deposited source linearly), and the device field copy
(planes and ring state, bit for bit)."""
from __future__ import annotations

import numpy as np
import pytest

import cases
from cases import quad_step
from genesis4_cuda_port_amd import api

pytestmark = pytest.mark.gpu

from test_gpu_parity import gpu_beam, gpu_field, noisy_field, relerr  # noqa: E402


@pytest.mark.parametrize("prec", [api.FP64, api.MIXED])
def test_beam_info(gpu_ctx, prec):
    c = cases.make_case(nslice=5, npart=128, ngrid=17)
    b = gpu_beam(gpu_ctx, c, prec)
    assert b.info() == (5, 5 * 128, prec)
    b.close()


def test_current_scales_the_source(gpu_ctx):
    """From a zero field, one field step is the source term alone, linear in
    the slice current: doubling the current of two slices doubles their
    planes and leaves the others as they were."""
    c = cases.make_case(nslice=4, npart=2048, ngrid=33)
    st = quad_step()
    z = np.zeros(c.nslice * c.ngrid ** 2)

    def step(current):
        b = gpu_beam(gpu_ctx, c)
        b.set_current(current)
        f = gpu_field(gpu_ctx, c, re=z, im=z)
        f.track(b, st, st.delz)
        re, im, _ = f.download()
        f.close()
        b.close()
        return (re + 1j * im).reshape(c.nslice, -1)

    cur = np.array(c.beam.current, dtype=float)
    a = step(cur)
    scale = np.array([1.0, 2.0, 1.0, 2.0])
    b = step(cur * scale)
    assert np.abs(a).max() > 0
    for s in range(c.nslice):
        assert relerr(b[s], scale[s] * a[s]) < 1e-13


@pytest.mark.parametrize("prec", [api.FP64, api.MIXED])
def test_field_copy_is_exact(gpu_ctx, prec):
    c = cases.make_case(nslice=6, npart=64, ngrid=21, seed_power=1e5)
    re, im = noisy_field(c, scale=0.5)
    src = gpu_field(gpu_ctx, c, prec, re=re, im=im, first=3)
    src.set_ring(3, 0.25)
    dst = gpu_field(gpu_ctx, c, prec)
    dst.copy_from(src)
    a, b = src.download(), dst.download()
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]) and a[2] == b[2] == 3
    assert dst.ring() == (3, 0.25)
    src.close()
    dst.close()
