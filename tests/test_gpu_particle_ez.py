"""GPU: the per-particle longitudinal field of the push (the UseParticleEz
kernels of BeamSolverCUDA.cpp:183-191,523-547) against the FP64 oracle, in
the single push and in the fused run; the per-slice energy change of the
collective effects (Beam.cpp:127-128) applied to the device beam."""
from __future__ import annotations

import numpy as np
import pytest

import _oracle
from _oracle import OField, dp, ip
import cases
from cases import GAMMA0, LAMBDA0, quad_step
from genesis4_cuda_port_amd import api
from genesis4_cuda_port_amd._abi import G4Error

pytestmark = pytest.mark.gpu

from test_gpu_parity import MIXED_TOL, gpu_beam, gpu_field, noisy_field, relerr  # noqa: E402


def _oracle_push(c, re, im, st, pez, first=2):
    of = OField(c.nslice, c.ngrid, c.gridmax, LAMBDA0, c.slicelength, 1, re, im, first=first)
    arr = (_oracle.OrcField * 1)(of.s)
    arr[0].re = dp(of.re)
    arr[0].im = dp(of.im)
    o = c.beam.copy()
    _oracle.lib().orc_beam_advance_ez(c.nslice, ip(o.offsets), dp(o.x), dp(o.y), dp(o.px),
                                      dp(o.py), dp(o.gamma), dp(o.theta), arr, 1, st, GAMMA0,
                                      st.delz, dp(pez))
    return o


@pytest.mark.parametrize("prec", [api.FP64, api.MIXED])
def test_push_particle_ez(gpu_ctx, prec):
    print(f"case push_particle_ez prec={prec}", flush=True)
    c = cases.make_case(nslice=5, npart=512, ngrid=33, seed_power=1e7)
    st = quad_step(autophase=0.3)
    re, im = noisy_field(c, scale=100.0)
    rng = np.random.default_rng(11)
    pez = rng.uniform(-2e-3, 2e-3, c.beam.gamma.size)
    b = gpu_beam(gpu_ctx, c, prec)
    f = gpu_field(gpu_ctx, c, prec, re=re, im=im, first=2)
    b.set_particle_ez(pez)
    b.advance([f], st, st.delz)
    g, t = b.download()[4:6]
    o = _oracle_push(c, re, im, st, pez)
    eg, et = relerr(g - GAMMA0, o.gamma - GAMMA0), relerr(t, o.theta)
    print(f"  gamma {eg:.3e} theta {et:.3e}", flush=True)
    if prec == api.FP64:
        assert eg < 1e-12 and et < 1e-13
    else:
        assert eg < MIXED_TOL and et < 1e-6
    # cleared: the push is the field-only one again
    b.upload(c.beam.x, c.beam.y, c.beam.px, c.beam.py, c.beam.gamma, c.beam.theta)
    b.set_particle_ez(None)
    b.advance([f], st, st.delz)
    g0 = b.download()[4]
    o0 = _oracle_push(c, re, im, st, np.zeros_like(pez))
    assert relerr(g0 - GAMMA0, o0.gamma - GAMMA0) < (1e-12 if prec == api.FP64 else MIXED_TOL)
    f.close()
    b.close()


def test_particle_ez_excludes_slice_loss(gpu_ctx):
    c = cases.make_case(nslice=2, npart=64, ngrid=17)
    st = quad_step()
    b = gpu_beam(gpu_ctx, c)
    f = gpu_field(gpu_ctx, c)
    b.set_particle_ez(np.zeros(c.beam.gamma.size))
    with pytest.raises(G4Error):
        b.advance([f], st, st.delz, np.zeros(c.nslice))
    with pytest.raises(ValueError):
        b.set_particle_ez(np.zeros(3))
    f.close()
    b.close()


@pytest.mark.parametrize("prec", [api.FP64, api.MIXED])
def test_run_particle_ez_zero_field(gpu_ctx, prec):
    """A fused run without radiation (no current, no seed): every particle's
    energy falls by its own ez times the distance, exactly in RK4."""
    print(f"case run_particle_ez prec={prec}", flush=True)
    c = cases.make_case(nslice=4, npart=1024, ngrid=31, nsteps=6, current=0.0, seed_power=0.0)
    rng = np.random.default_rng(4)
    pez = rng.uniform(-1e-3, 1e-3, c.beam.gamma.size)
    b = gpu_beam(gpu_ctx, c, prec)
    f = gpu_field(gpu_ctx, c, prec)
    b.set_particle_ez(pez)
    run = api.Run(gpu_ctx, b, f, c.steps, LAMBDA0, c.sample, c.timerun, False, 1)
    run.advance(len(c.steps))
    g = b.download()[4]
    run.close()
    f.close()
    b.close()
    zlen = sum(s.delz for s in c.steps)
    want = c.beam.gamma - pez * zlen
    err = np.abs(g - want).max()
    print(f"  max |dgamma error| {err:.3e}", flush=True)
    assert err < (1e-9 if prec == api.FP64 else 2e-3)


@pytest.mark.parametrize("prec", [api.FP64, api.MIXED])
def test_add_energy_per_slice(gpu_ctx, prec):
    """Every particle of slice s gains dgamma[s]; nothing else moves.
    Includes an empty slice."""
    print(f"case add_energy prec={prec}", flush=True)
    c = cases.make_case(nslice=5, npart=256, ngrid=17)
    off = c.beam.offsets.copy()
    off[2] = off[3]  # the third slice is empty, the second takes its particles
    dg = np.array([1e-3, -2.5e-3, 7.0, 0.0, 0.125])
    b = api.Beam(gpu_ctx, prec, off, np.full(c.nslice, 1000.0), GAMMA0)
    b.upload(c.beam.x, c.beam.y, c.beam.px, c.beam.py, c.beam.gamma, c.beam.theta)
    before = b.download()
    b.add_energy(dg)
    after = b.download()
    want = np.repeat(dg, np.diff(off))
    d = after[4] - before[4]
    if prec == api.FP64:
        assert np.array_equal(after[4], before[4] + want)
    else:
        assert np.abs(d - want).max() < 4e-6
    for k in (0, 1, 2, 3, 5):
        assert np.array_equal(after[k], before[k])
    b.close()
