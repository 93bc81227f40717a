"""GPU parity: the HIP path (through the C ABI) against the FP64 CPU oracle.

Bars (stated per assertion):
  * particle-to-cell indices, slippage shifts and particle counts: bit-exact;
  * FP64 mode: 1e-10 relative (normwise over slices and records) on slice
    power and bunching, and on the operator outputs;
  * MIXED mode: 1e-3 relative on slice power and bunching against the FP64
    oracle.
"""
from __future__ import annotations

import numpy as np
import pytest

import _oracle
from _oracle import OField, dp, ip
import cases
from cases import GAMMA0, LAMBDA0, quad_step
from genesis4_cuda_port_amd import api
from genesis4_cuda_port_amd._abi import FIELD_RAW, BEAM_STATS

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-10
MIXED_TOL = 1e-3


def relerr(a, b):
    """Normwise relative error ||a-b|| / ||b|| (complex-aware)."""
    a = np.asarray(a)
    b = np.asarray(b)
    den = np.linalg.norm(b.ravel())
    if den == 0:
        return np.linalg.norm(a.ravel())
    return np.linalg.norm((a - b).ravel()) / den


def gpu_beam(ctx, c, prec=api.FP64):
    b = api.Beam(ctx, prec, c.beam.offsets, c.beam.current, GAMMA0)
    b.upload(c.beam.x, c.beam.y, c.beam.px, c.beam.py, c.beam.gamma, c.beam.theta)
    return b


def gpu_field(ctx, c, prec=api.FP64, kind=api.ADI, harm=1, re=None, im=None, first=0):
    f = api.Field(ctx, prec, c.nslice, c.ngrid, c.gridmax, LAMBDA0, c.slicelength, harm)
    f.upload(c.seed_re if re is None else re, c.seed_im if im is None else im, first)
    f.init_solver(kind)
    return f


def noisy_field(c, seed=7, scale=1.0):
    rng = np.random.default_rng(seed)
    n = c.nslice * c.ngrid * c.ngrid
    re = c.seed_re + scale * 1e3 * rng.standard_normal(n)
    im = c.seed_im + scale * 1e3 * rng.standard_normal(n)
    return re, im


# K dz^2 of the half step, bounded for gamma >= gamma0/2: 5e-4 at qf = 1.5
# (the short series of the plane maps), 0.034 / 0.10 at qf = 100 / -300 (the
# long series), 1.7 at qf = 5000 (the trigonometric form)
TRANSVERSE_STEPS = {
    "undulator": dict(),
    "quad_focus": dict(qf=1.5),
    "quad_defocus": dict(qf=-1.5),
    "quad_medium": dict(qf=100.0),
    "quad_medium_defocus": dict(qf=-300.0),
    "quad_strong": dict(qf=5000.0),
    "quad_strong_defocus": dict(qf=-5000.0),
    "drift": dict(aw=0.0, in_und=False),
    "offsets": dict(qf=1.5, dqx=1e-5, dqy=-2e-5, ax=3e-6, ay=-4e-6),
    "corrector": dict(cx=1e-6, cy=-2e-6),
    "chicane": dict(chicane=(0.01, 0.02, 0.1, 0.5)),
}


@pytest.mark.parametrize("name", list(TRANSVERSE_STEPS))
@pytest.mark.parametrize("last", [False, True])
def test_track_transverse_fp64(gpu_ctx, name, last):
    print(f"case track_transverse {name} last={last}", flush=True)
    c = cases.make_case(nslice=4, npart=512)
    st = quad_step(**TRANSVERSE_STEPS[name])
    b = gpu_beam(gpu_ctx, c)
    b.track_transverse(st, st.delz * 0.5, last)
    x, y, px, py, g, t = b.download()
    o = c.beam.copy()
    _oracle.lib().orc_track_transverse(o.x.size, dp(o.x), dp(o.y), dp(o.px), dp(o.py),
                                       dp(o.gamma), st, GAMMA0, st.delz * 0.5, int(last))
    for a, r in ((x, o.x), (y, o.y), (px, o.px), (py, o.py)):
        assert relerr(a, r) < 1e-13
    assert np.array_equal(g, o.gamma) and np.array_equal(t, o.theta)


@pytest.mark.parametrize("name", ["undulator", "quad_focus", "quad_medium_defocus",
                                  "quad_strong", "chicane"])
def test_track_transverse_mixed(gpu_ctx, name):
    """MIXED transverse maps (FP32 coordinates; short and long series, the
    trigonometric form, the chicane) against the FP64 oracle: normwise 1e-6,
    a few FP32 roundings of each coordinate."""
    print(f"case track_transverse_mixed {name}", flush=True)
    c = cases.make_case(nslice=4, npart=512)
    st = quad_step(**TRANSVERSE_STEPS[name])
    b = gpu_beam(gpu_ctx, c, api.MIXED)
    b.track_transverse(st, st.delz * 0.5, False)
    x, y, px, py, g, t = b.download()
    o = c.beam.copy()
    _oracle.lib().orc_track_transverse(o.x.size, dp(o.x), dp(o.y), dp(o.px), dp(o.py),
                                       dp(o.gamma), st, GAMMA0, st.delz * 0.5, 0)
    for a, r in ((x, o.x), (y, o.y), (px, o.px), (py, o.py)):
        assert relerr(a, r) < 1e-6


def test_r56_fp64(gpu_ctx):
    c = cases.make_case(nslice=4, npart=256)
    st = quad_step(chicane=(0.02, 0.05, 0.2, 1.0))
    b = gpu_beam(gpu_ctx, c)
    b.apply_r56(st, LAMBDA0)
    t = b.download()[5]
    o = c.beam.copy()
    _oracle.lib().orc_apply_r56(o.x.size, dp(o.gamma), dp(o.theta), st, GAMMA0, LAMBDA0)
    assert relerr(t, o.theta) < 1e-14


@pytest.mark.parametrize("nfield", [1, 2])
@pytest.mark.parametrize("eloss", [False, True])
def test_beam_advance_fp64(gpu_ctx, nfield, eloss):
    print(f"case beam_advance nfield={nfield} eloss={eloss}", flush=True)
    c = cases.make_case(nslice=6, npart=512, ngrid=33, seed_power=1e7)
    st = quad_step(autophase=0.3, gradx=2.0, ax=1e-6)
    re, im = noisy_field(c, scale=100.0)
    fields = [gpu_field(gpu_ctx, c, re=re, im=im, first=2)]
    ofields = [OField(c.nslice, c.ngrid, c.gridmax, LAMBDA0, c.slicelength, 1, re, im, first=2)]
    if nfield == 2:
        re3, im3 = noisy_field(c, seed=9, scale=30.0)
        fields.append(gpu_field(gpu_ctx, c, harm=3, re=re3, im=im3, first=2))
        ofields.append(OField(c.nslice, c.ngrid, c.gridmax, LAMBDA0, c.slicelength, 3, re3, im3,
                              first=2))
    el = (np.linspace(-1e-3, 1e-3, c.nslice) if eloss else None)
    b = gpu_beam(gpu_ctx, c)
    b.advance(fields, st, st.delz, el)
    g, t = b.download()[4:6]
    o = c.beam.copy()
    arr = (_oracle.OrcField * nfield)(*[f.s for f in ofields])
    for k, f in enumerate(ofields):
        arr[k].re = dp(f.re)
        arr[k].im = dp(f.im)
    _oracle.lib().orc_beam_advance(c.nslice, ip(o.offsets), dp(o.x), dp(o.y), dp(o.px), dp(o.py),
                                   dp(o.gamma), dp(o.theta), arr, nfield, st, GAMMA0, st.delz,
                                   None if el is None else dp(el))
    assert relerr(g - GAMMA0, o.gamma - GAMMA0) < 1e-12
    assert relerr(t, o.theta) < 1e-13


@pytest.mark.parametrize("prec,nfield", [(api.FP64, 8), (api.MIXED, 3)])
def test_beam_advance_many_fields(gpu_ctx, prec, nfield):
    """The general multi-field push up to its limit of 8 fields (harmonics
    1..nfield), against the oracle."""
    print(f"case beam_advance_many prec={prec} nfield={nfield}", flush=True)
    c = cases.make_case(nslice=4, npart=512, ngrid=33, seed_power=1e7)
    st = quad_step(autophase=0.1)
    fields, ofields = [], []
    for h in range(1, nfield + 1):
        re, im = noisy_field(c, seed=20 + h, scale=100.0 / h)
        fields.append(gpu_field(gpu_ctx, c, prec, harm=h, re=re, im=im, first=1))
        ofields.append(OField(c.nslice, c.ngrid, c.gridmax, LAMBDA0, c.slicelength, h, re, im,
                              first=1))
    b = gpu_beam(gpu_ctx, c, prec)
    b.advance(fields, st, st.delz)
    g, t = b.download()[4:6]
    o = c.beam.copy()
    arr = (_oracle.OrcField * nfield)(*[f.s for f in ofields])
    for k, f in enumerate(ofields):
        arr[k].re = dp(f.re)
        arr[k].im = dp(f.im)
    _oracle.lib().orc_beam_advance(c.nslice, ip(o.offsets), dp(o.x), dp(o.y), dp(o.px), dp(o.py),
                                   dp(o.gamma), dp(o.theta), arr, nfield, st, GAMMA0, st.delz,
                                   None)
    eg, et = relerr(g - GAMMA0, o.gamma - GAMMA0), relerr(t, o.theta)
    print(f"  gamma {eg:.3e} theta {et:.3e}", flush=True)
    if prec == api.FP64:
        assert eg < 1e-12 and et < 1e-13
    else:
        assert eg < MIXED_TOL and et < 1e-6
    for f in fields:
        f.close()
    b.close()


def test_cell_index_bitexact(gpu_ctx):
    """Indices of random and adversarial positions (grid nodes, +-gridmax,
    one ulp either side, NaN) must equal the oracle's exactly."""
    c = cases.make_case(nslice=2, npart=4096, ngrid=31)
    rng = np.random.default_rng(3)
    n = c.beam.x.size
    gm = c.gridmax
    dg = 2 * gm / (c.ngrid - 1)
    nodes = -gm + dg * rng.integers(0, c.ngrid, n)
    x = np.where(rng.random(n) < 0.5, rng.uniform(-1.1 * gm, 1.1 * gm, n), nodes)
    y = np.where(rng.random(n) < 0.5, rng.uniform(-1.1 * gm, 1.1 * gm, n), nodes[::-1])
    special = np.array([gm, -gm, np.nextafter(gm, 0), np.nextafter(-gm, 0), 0.0,
                        np.nextafter(gm - dg, gm), gm - dg, np.nextafter(gm - dg, 0), np.nan])
    x[:special.size] = special
    y[:special.size] = special[::-1]
    bm = c.beam.copy()
    bm.x[:] = x
    bm.y[:] = y
    b = api.Beam(gpu_ctx, api.FP64, bm.offsets, bm.current, GAMMA0)
    b.upload(bm.x, bm.y, bm.px, bm.py, bm.gamma, bm.theta)
    f = gpu_field(gpu_ctx, c)
    got = b.cell_index(f)
    ref = np.empty(n, dtype=np.int32)
    _oracle.lib().orc_cell_index(n, dp(bm.x), dp(bm.y), gm, f.dgrid, c.ngrid, ip(ref))
    assert np.array_equal(got, ref)
    assert (ref == -1).any() and (ref >= 0).any()


@pytest.mark.parametrize("kind", [api.ADI, api.FFT])
@pytest.mark.parametrize("ngrid", [32, 33])
def test_field_track_fp64(gpu_ctx, kind, ngrid):
    """Field::track = deposition + one solver step; two steps so both the
    first-step (global atomic) and the windowed LDS deposition run."""
    print(f"case field_track kind={kind} ngrid={ngrid}", flush=True)
    c = cases.make_case(nslice=6, npart=1024, ngrid=ngrid, seed_power=1e6)
    st = quad_step()
    re, im = noisy_field(c, scale=0.1)
    f = gpu_field(gpu_ctx, c, kind=kind, re=re, im=im, first=1)
    b = gpu_beam(gpu_ctx, c)
    of = OField(c.nslice, c.ngrid, c.gridmax, LAMBDA0, c.slicelength, 1, re, im, first=1)
    o = c.beam
    plane = c.ngrid * c.ngrid
    for it in range(2):
        f.track(b, st, st.delz)
        sre = np.zeros(c.nslice * plane)
        sim = np.zeros_like(sre)
        on = _oracle.lib().orc_deposit(c.nslice, ip(o.offsets), dp(o.x), dp(o.y), dp(o.gamma),
                                       dp(o.theta), of.ptr, st, dp(o.current), st.delz, dp(sre),
                                       dp(sim))
        assert on == 1
        if kind == api.ADI:
            _oracle.lib().orc_adi_advance(of.ptr, dp(sre), dp(sim), st.delz)
        else:
            _oracle.lib().orc_fft_advance(of.ptr, dp(sre), dp(sim), st.delz, 0, 1.0, 1.0, 1.0)
        gre, gim, first = f.download()
        assert first == 1
        e = relerr(gre + 1j * gim, of.re + 1j * of.im)
        print(f"  step {it}: rel err {e:.3e}", flush=True)
        assert e < 1e-12


@pytest.mark.parametrize("prec,ngrid", [(api.FP64, g) for g in (4, 8, 16, 64, 128, 256, 512, 1024,
                                                                   3, 5, 15, 33, 48, 96, 151, 255,
                                                                   500, 600)]
                         + [(api.MIXED, g) for g in (16, 64, 256, 1024, 33, 151, 255)])
def test_fft_sizes(gpu_ctx, prec, ngrid):
    """Spectral step at every transform shape: one or several lines per wave
    (G < 64), one to sixteen points per lane (G >= 64); sizes that are not
    powers of two up to 512 through the convolution by M-point transforms
    (M = 8 ... 1024), larger ones (600) through the direct DFT."""
    print(f"case fft_sizes prec={prec} ngrid={ngrid}", flush=True)
    ns = 1 if ngrid == 600 else 3  # the oracle's direct DFT at 600 takes seconds per plane
    first = 2 % ns
    c = cases.make_case(nslice=ns, npart=256, ngrid=ngrid, seed_power=1e6)
    st = quad_step()
    re, im = noisy_field(c, scale=0.1)
    f = gpu_field(gpu_ctx, c, prec, kind=api.FFT, re=re, im=im, first=first)
    b = gpu_beam(gpu_ctx, c, prec)
    of = OField(c.nslice, c.ngrid, c.gridmax, LAMBDA0, c.slicelength, 1, re, im, first=first)
    o = c.beam
    tol = 1e-12 if prec == api.FP64 else 2e-5
    for it in range(2):
        f.track(b, st, st.delz)
        sre = np.zeros(c.nslice * c.ngrid * c.ngrid)
        sim = np.zeros_like(sre)
        _oracle.lib().orc_deposit(c.nslice, ip(o.offsets), dp(o.x), dp(o.y), dp(o.gamma),
                                  dp(o.theta), of.ptr, st, dp(o.current), st.delz, dp(sre),
                                  dp(sim))
        _oracle.lib().orc_fft_advance(of.ptr, dp(sre), dp(sim), st.delz, 0, 1.0, 1.0, 1.0)
        gre, gim, ring = f.download()
        assert ring == first
        e = relerr(gre + 1j * gim, of.re + 1j * of.im)
        print(f"  step {it}: rel err {e:.3e}", flush=True)
        assert e < tol
    fraw = f.diag_raw()
    ofraw = np.zeros_like(fraw)
    _oracle.lib().orc_diag_field_raw(of.ptr, dp(ofraw))
    # field sums weight the far cells by dx^2, where the round-off of two
    # different FFT factorisations is largest relative to the field: the FP64
    # bar of 1e-10 applies here, not the 1e-12 of the field itself
    assert relerr(fraw, ofraw) < (1e-10 if prec == api.FP64 else 1e-4)


@pytest.mark.parametrize("ngrid", [32, 128, 256, 33])
def test_fft_filter_fp64(gpu_ctx, ngrid):
    c = cases.make_case(nslice=3, npart=1024, ngrid=ngrid, seed_power=1e6)
    st = quad_step()
    f = gpu_field(gpu_ctx, c, kind=api.FFT)
    f.init_solver(api.FFT, source_filter=True, xcut=0.3, ycut=0.4, sigmoid=0.1)
    b = gpu_beam(gpu_ctx, c)
    of = OField(c.nslice, c.ngrid, c.gridmax, LAMBDA0, c.slicelength, 1, c.seed_re, c.seed_im)
    f.track(b, st, st.delz)
    o = c.beam
    sre = np.zeros(c.nslice * c.ngrid ** 2)
    sim = np.zeros_like(sre)
    _oracle.lib().orc_deposit(c.nslice, ip(o.offsets), dp(o.x), dp(o.y), dp(o.gamma), dp(o.theta),
                              of.ptr, st, dp(o.current), st.delz, dp(sre), dp(sim))
    _oracle.lib().orc_fft_advance(of.ptr, dp(sre), dp(sim), st.delz, 1, 0.3, 0.4, 0.1)
    gre, gim, _ = f.download()
    assert relerr(gre + 1j * gim, of.re + 1j * of.im) < 1e-12


@pytest.mark.parametrize("ngrid", [256, 128, 33])
def test_fft_filter_mixed(gpu_ctx, ngrid):
    """The filtered source in MIXED (FP64 work planes, FP32 field) against the
    FP64 oracle: bar 1e-5 relative (FP32 round-off of one step is ~2e-7)."""
    print(f"case fft filter mixed ngrid={ngrid}", flush=True)
    c = cases.make_case(nslice=3, npart=1024, ngrid=ngrid, seed_power=1e6)
    st = quad_step()
    f = gpu_field(gpu_ctx, c, api.MIXED, kind=api.FFT)
    f.init_solver(api.FFT, source_filter=True, xcut=0.3, ycut=0.4, sigmoid=0.1)
    b = gpu_beam(gpu_ctx, c, api.MIXED)
    of = OField(c.nslice, c.ngrid, c.gridmax, LAMBDA0, c.slicelength, 1, c.seed_re, c.seed_im)
    f.track(b, st, st.delz)
    o = c.beam
    sre = np.zeros(c.nslice * c.ngrid ** 2)
    sim = np.zeros_like(sre)
    _oracle.lib().orc_deposit(c.nslice, ip(o.offsets), dp(o.x), dp(o.y), dp(o.gamma), dp(o.theta),
                              of.ptr, st, dp(o.current), st.delz, dp(sre), dp(sim))
    _oracle.lib().orc_fft_advance(of.ptr, dp(sre), dp(sim), st.delz, 1, 0.3, 0.4, 0.1)
    gre, gim, _ = f.download()
    err = relerr(gre + 1j * gim, of.re + 1j * of.im)
    print(f"  field {err:.3e}", flush=True)
    assert err < 1e-5


@pytest.mark.parametrize("slips", [[1.0] * 5, [0.5, 0.5, 0.5, 2.5, -1.0, -1.0], [3.0]])
@pytest.mark.parametrize("periodic", [False, True])
def test_slippage_bitexact(gpu_ctx, slips, periodic):
    c = cases.make_case(nslice=7, npart=16, ngrid=9, seed_power=1e6)
    f = gpu_field(gpu_ctx, c, first=3)
    of = OField(c.nslice, c.ngrid, c.gridmax, LAMBDA0, c.slicelength, 1, c.seed_re, c.seed_im,
                first=3)
    for s in slips:
        n = f.apply_slippage(s, 1.0, True, periodic)
        m = _oracle.lib().orc_apply_slippage(of.ptr, s, 1.0, 1, int(periodic))
        assert n == m
        re, im, first = f.download()
        assert first == of.first
        assert np.array_equal(re, of.re) and np.array_equal(im, of.im)
        assert f.ring()[1] == of.s.accuslip


@pytest.mark.parametrize("nharm", [3, 8])
@pytest.mark.parametrize("prec", [api.FP64, api.MIXED])
def test_diag_raw(gpu_ctx, prec, nharm):
    c = cases.make_case(nslice=5, npart=1024, ngrid=33, seed_power=1e6)
    b = gpu_beam(gpu_ctx, c, prec)
    re, im = noisy_field(c, scale=0.01)
    f = gpu_field(gpu_ctx, c, prec, re=re, im=im, first=2)
    of = OField(c.nslice, c.ngrid, c.gridmax, LAMBDA0, c.slicelength, 1, re, im, first=2)
    braw = b.diag_raw(nharm)
    o = c.beam
    oraw = np.zeros_like(braw)
    _oracle.lib().orc_diag_beam_raw(c.nslice, ip(o.offsets), dp(o.x), dp(o.y), dp(o.px),
                                    dp(o.py), dp(o.gamma), dp(o.theta), nharm, GAMMA0, dp(oraw))
    fraw = f.diag_raw()
    ofraw = np.zeros_like(fraw)
    _oracle.lib().orc_diag_field_raw(of.ptr, dp(ofraw))
    tol = 1e-12 if prec == api.FP64 else 1e-5
    assert np.array_equal(braw[:, 12], oraw[:, 12])  # particle counts: exact
    for k in range(BEAM_STATS):
        if k == 12:
            continue
        assert relerr(braw[:, k], oraw[:, k]) < (tol if k > 1 else max(tol, 1e-6)), k
    assert relerr(braw[:, BEAM_STATS:], oraw[:, BEAM_STATS:]) < tol * 10
    assert relerr(fraw, ofraw) < tol


def _run_pair(ctx, c, kind, prec, nharm=2):
    steps = c.steps
    b = gpu_beam(ctx, c, prec)
    f = gpu_field(ctx, c, prec, kind=kind)
    run = api.Run(ctx, b, f, steps, LAMBDA0, c.sample, c.timerun, False, nharm)
    run.advance(len(steps))
    braw, fraw = run.download_diag()
    of = OField(c.nslice, c.ngrid, c.gridmax, LAMBDA0, c.slicelength, 1, c.seed_re, c.seed_im)
    ob, obraw, ofraw = _oracle.run(c.beam, of, steps, gamma0=GAMMA0, lambda0=LAMBDA0, solver=kind,
                                   sample=c.sample, timerun=c.timerun, nharm=nharm)
    return (b, f, braw, fraw), (ob, of, obraw, ofraw)


def _power_bunching(braw, fraw, c, f):
    n = braw.shape[0]
    P = np.empty((n, c.nslice))
    B = np.empty((n, c.nslice))
    for r in range(n):
        fq, _ = api.finalize_field(fraw[r], f.dgrid, f.xlambda, f.slicelength)
        bq, _ = api.finalize_beam(braw[r], 2, c.beam.current, GAMMA0)
        P[r] = fq[:, 0]
        B[r] = bq[:, 25]
    return P, B


@pytest.mark.parametrize("kind", [api.ADI, api.FFT])
def test_run_fp64(gpu_ctx, kind):
    print(f"case run_fp64 kind={kind}", flush=True)
    c = cases.make_case(nslice=16, npart=512, ngrid=32 if kind == api.FFT else 31, nsteps=30,
                        seed_power=0.0)
    (b, f, braw, fraw), (ob, of, obraw, ofraw) = _run_pair(gpu_ctx, c, kind, api.FP64)
    assert braw.shape == obraw.shape
    P, B = _power_bunching(braw, fraw, c, f)
    oP, oB = _power_bunching(obraw, ofraw, c, f)
    ep, eb = relerr(P[1:], oP[1:]), relerr(B, oB)
    print(f"  power rel err {ep:.3e}  bunching rel err {eb:.3e}", flush=True)
    assert ep < FP64_TOL
    assert eb < FP64_TOL
    assert f.ring()[0] == of.first
    x, y, px, py, g, t = b.download()
    assert relerr(g - GAMMA0, ob.gamma - GAMMA0) < 1e-10
    assert relerr(t, ob.theta) < 1e-12


@pytest.mark.parametrize("kind", [api.ADI, api.FFT])
def test_run_mixed(gpu_ctx, kind):
    print(f"case run_mixed kind={kind}", flush=True)
    c = cases.make_case(nslice=16, npart=512, ngrid=32 if kind == api.FFT else 31, nsteps=30,
                        seed_power=1e3)
    (b, f, braw, fraw), (ob, of, obraw, ofraw) = _run_pair(gpu_ctx, c, kind, api.MIXED)
    P, B = _power_bunching(braw, fraw, c, f)
    oP, oB = _power_bunching(obraw, ofraw, c, f)
    ep, eb = relerr(P, oP), relerr(B, oB)
    print(f"  power rel err {ep:.3e}  bunching rel err {eb:.3e}", flush=True)
    assert ep < MIXED_TOL
    assert eb < MIXED_TOL
