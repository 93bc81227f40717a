"""GPU: short-range space charge computed on the device (g4h_beam_init_efield;
the reference's EFieldSolver::shortRange on the host, BeamSolverCUDA.cpp:
523-547) against the FP64 oracle's orc_sc_field, in the single push, in the
fused run (both solvers, both precisions) and in a replayed graph.

Bars: FP64 push 1e-12 relative on gamma - gamma0 and theta, per-slice RMS
field 1e-9 (MIXED 1e-3); runs 1e-10 (FP64) and 1e-3 (MIXED) relative on slice power and
bunching, as for the runs without space charge.  The device sums the mode
sources in LDS in no fixed order, so agreement is to round-off."""
from __future__ import annotations

import numpy as np
import pytest

import _oracle
from _oracle import OField, dp, ip
import cases
from cases import AW, GAMMA0, LAMBDA0, quad_step
from genesis4_cuda_port_amd import api
from genesis4_cuda_port_amd._abi import G4Error

pytestmark = pytest.mark.gpu

from test_gpu_parity import (FP64_TOL, MIXED_TOL, _power_bunching, gpu_beam, gpu_field,  # noqa: E402
                             noisy_field, relerr)

# rmax, shells, longitudinal harmonics, azimuthal modes
EF = (2.0, 40, 2, 1)
KS = 2 * np.pi / LAMBDA0


def _bunched_case(nslice=6, npart=1024, ngrid=33, nsteps=12, seed_power=1e6):
    """A shot-noise beam whose phases are pulled towards 0 (bunching ~0.3),
    so that the space-charge field is well above round-off."""
    c = cases.make_case(nslice=nslice, npart=npart, ngrid=ngrid, nsteps=nsteps,
                        seed_power=seed_power)
    c.beam.theta = c.beam.theta - 0.3 * np.sin(c.beam.theta)
    return c


@pytest.mark.parametrize("prec", [api.FP64, api.MIXED])
def test_push_with_space_charge(gpu_ctx, prec):
    print(f"case sc push prec={prec}", flush=True)
    c = _bunched_case()
    st = quad_step(autophase=0.2)
    re, im = noisy_field(c, scale=10.0)
    eloss = np.linspace(-1e-4, 1e-4, c.nslice)
    b = gpu_beam(gpu_ctx, c, prec)
    f = gpu_field(gpu_ctx, c, prec, re=re, im=im, first=2)
    b.init_efield(*EF, LAMBDA0)
    b.advance([f], st, st.delz, eloss)
    g, t = b.download()[4:6]
    rms = b.sc_field()
    gz2 = GAMMA0 ** 2 / (1 + st.aw ** 2)
    ez, orms = _oracle.sc_field(c.beam, KS, gz2, *EF)
    assert np.all(orms > 0)
    assert np.max(np.abs(ez)) > 1e-4  # the field matters at this bunching
    ez = ez + np.repeat(eloss, np.diff(c.beam.offsets))
    of = OField(c.nslice, c.ngrid, c.gridmax, LAMBDA0, c.slicelength, 1, re, im, first=2)
    arr = (_oracle.OrcField * 1)(of.s)
    arr[0].re = dp(of.re)
    arr[0].im = dp(of.im)
    o = c.beam.copy()
    _oracle.lib().orc_beam_advance_ez(c.nslice, ip(o.offsets), dp(o.x), dp(o.y), dp(o.px),
                                      dp(o.py), dp(o.gamma), dp(o.theta), arr, 1, st, GAMMA0,
                                      st.delz, dp(ez))
    eg, et, em = relerr(g - GAMMA0, o.gamma - GAMMA0), relerr(t, o.theta), relerr(rms, orms)
    print(f"  gamma {eg:.3e} theta {et:.3e} slice RMS field {em:.3e}", flush=True)
    if prec == api.FP64:
        assert eg < 1e-12 and et < 1e-13 and em < 1e-9
    else:
        assert eg < MIXED_TOL and et < 1e-6 and em < 1e-3


@pytest.mark.parametrize("prec", [api.FP64, api.MIXED])
@pytest.mark.parametrize("kind", [api.ADI, api.FFT])
def test_run_with_space_charge(gpu_ctx, prec, kind):
    print(f"case sc run prec={prec} kind={kind}", flush=True)
    c = _bunched_case(nslice=8, npart=512, ngrid=31 if kind == api.ADI else 32, nsteps=20)
    b = gpu_beam(gpu_ctx, c, prec)
    f = gpu_field(gpu_ctx, c, prec, kind=kind)
    b.init_efield(*EF, LAMBDA0)
    run = api.Run(gpu_ctx, b, f, c.steps, LAMBDA0, c.sample, c.timerun, False, 2)
    run.advance(len(c.steps))
    braw, fraw = run.download_diag()
    of = OField(c.nslice, c.ngrid, c.gridmax, LAMBDA0, c.slicelength, 1, c.seed_re, c.seed_im)
    _, obraw, ofraw = _oracle.run(c.beam, of, c.steps, gamma0=GAMMA0, lambda0=LAMBDA0,
                                  solver=kind, sample=c.sample, timerun=c.timerun, nharm=2,
                                  efield=EF)
    # without space charge the run differs well above the bar
    of0 = OField(c.nslice, c.ngrid, c.gridmax, LAMBDA0, c.slicelength, 1, c.seed_re, c.seed_im)
    _, nbraw, nfraw = _oracle.run(c.beam, of0, c.steps, gamma0=GAMMA0, lambda0=LAMBDA0,
                                  solver=kind, sample=c.sample, timerun=c.timerun, nharm=2)
    P, B = _power_bunching(braw, fraw, c, f)
    oP, oB = _power_bunching(obraw, ofraw, c, f)
    nP, nB = _power_bunching(nbraw, nfraw, c, f)
    ep, eb = relerr(P[1:], oP[1:]), relerr(B, oB)
    dsc = relerr(B, nB)
    print(f"  power {ep:.3e} bunching {eb:.3e} (space charge changes bunching by {dsc:.2e})",
          flush=True)
    tol = FP64_TOL if prec == api.FP64 else MIXED_TOL
    assert ep < tol and eb < tol
    # the precision resolves the space-charge effect
    assert dsc > 100 * eb
    assert f.ring()[0] == of.first


def test_graph_replay_with_space_charge():
    print("case sc graph replay", flush=True)
    c = _bunched_case(nslice=8, npart=512, ngrid=31, nsteps=10)
    with api.Context(0) as ctx:
        b = gpu_beam(ctx, c, api.FP64)
        f = gpu_field(ctx, c, api.FP64, kind=api.ADI)
        b.init_efield(*EF, LAMBDA0)
        run = api.Run(ctx, b, f, c.steps, LAMBDA0, c.sample, c.timerun, False, 1)
        run.advance(len(c.steps))
        eager = run.download_diag()
        run.capture()
        for _ in range(2):
            b.upload(c.beam.x, c.beam.y, c.beam.px, c.beam.py, c.beam.gamma, c.beam.theta)
            f.upload(c.seed_re, c.seed_im, 0)
            run.replay()
            braw, fraw = run.download_diag()
            assert relerr(braw, eager[0]) < 1e-12 and relerr(fraw, eager[1]) < 1e-12
            b.diag_raw(1)  # eager work between replays
            ctx.synchronize()
        # new settings after the capture: the replay is refused
        b.init_efield(3.0, 40, 1, 0, LAMBDA0)
        b.upload(c.beam.x, c.beam.y, c.beam.px, c.beam.py, c.beam.gamma, c.beam.theta)
        f.upload(c.seed_re, c.seed_im, 0)
        with pytest.raises(G4Error):
            run.replay()
        run.close()
        f.close()
        b.close()


def test_efield_settings_checked(gpu_ctx):
    print("case sc settings", flush=True)
    c = cases.make_case(nslice=2, npart=64, ngrid=17)
    b = gpu_beam(gpu_ctx, c)
    with pytest.raises(G4Error):
        b.init_efield(2.0, 40, 1, 0, LAMBDA0, longrange=True)
    with pytest.raises(G4Error):
        b.init_efield(0.0, 40, 1, 0, LAMBDA0)
    with pytest.raises(G4Error):
        b.init_efield(2.0, 1000, 4, 2, LAMBDA0)  # 20 modes x 1000 shells: beyond the LDS
    b.init_efield(2.0, 40, 0, 0, LAMBDA0)  # off
    assert np.array_equal(b.sc_field(), np.zeros(c.nslice))
    b.close()
