"""GPU: kernel variants and the autotuner.

The three single-chain ADI variants (workgroup per plane; wave per 64 lines
with 8- or 16-element load chunks) must give bit-identical fields and field
sums; each of them and the segmented variant 3 (16 waves per 64 lines, every
element in registers) must match the FP64 oracle with the source on; the
lane variants must cover grids the workgroup variant cannot hold in LDS
(G > 360); an unpinned context runs the segmented sweeps for G <= 256
(equal bit for bit to a run pinned to 3, and to the single-chain runs within
round-off, FP64 1e-13 and MIXED 1e-6 on the fields and sums), and on larger
grids the tuner must settle on one of 0-2 after its trial launches without
changing results."""
from __future__ import annotations

import numpy as np
import pytest

import _oracle
from _oracle import OField, dp, ip
import cases
from cases import LAMBDA0, quad_step
from genesis4_cuda_port_amd import api
from genesis4_cuda_port_amd._abi import TUNE_ADI

pytestmark = pytest.mark.gpu

from test_gpu_parity import gpu_beam, gpu_field, noisy_field, relerr  # noqa: E402


def _run_records(prec, variant, c, re, im, nsteps):
    with api.Context(0) as ctx:
        ctx.set_variant(TUNE_ADI, variant)
        b = gpu_beam(ctx, c, prec)
        f = gpu_field(ctx, c, prec, kind=api.ADI, re=re, im=im, first=2)
        run = api.Run(ctx, b, f, c.steps[:nsteps], LAMBDA0, c.sample, True, False, 1)
        assert run.advance(nsteps) == nsteps
        braw, fraw = run.download_diag()
        gre, gim, first = f.download()
        run.close()
        f.close()
        b.close()
    return gre, gim, fraw


@pytest.mark.parametrize("prec", [api.FP64, api.MIXED])
@pytest.mark.parametrize("ngrid", [31, 64, 65, 151])
def test_adi_variants_bitexact(prec, ngrid):
    """No beam current, so no deposition (whose LDS atomics may sum in any
    order): the field steps and the fused field sums of every record must be
    identical bit for bit between the variants."""
    print(f"case adi_variants prec={prec} ngrid={ngrid}", flush=True)
    c = cases.make_case(nslice=5, npart=64, ngrid=ngrid, seed_power=1e6, current=0.0, nsteps=4)
    re, im = noisy_field(c, scale=0.1)
    a = _run_records(prec, 0, c, re, im, 4)
    for v in (1, 2):
        b = _run_records(prec, v, c, re, im, 4)
        for x, y in zip(a, b):
            assert np.array_equal(x, y)
    assert np.abs(a[0]).max() > 0


@pytest.mark.parametrize("variant,ngrid", [(0, 33), (1, 33), (1, 64), (1, 65), (1, 400),
                                            (2, 33), (2, 65), (2, 400), (1, 1024), (2, 1024),
                                            (3, 9), (3, 17), (3, 33), (3, 64), (3, 65), (3, 101),
                                            (3, 151), (3, 255), (3, 256), (3, 400)])
def test_adi_variant_oracle_fp64(variant, ngrid):
    """Deposition + ADI step with the source on, each variant pinned."""
    print(f"case adi_variant_oracle variant={variant} ngrid={ngrid}", flush=True)
    c = cases.make_case(nslice=3, npart=512, ngrid=ngrid, seed_power=1e6)
    st = quad_step()
    re, im = noisy_field(c, scale=0.1)
    of = OField(c.nslice, c.ngrid, c.gridmax, LAMBDA0, c.slicelength, 1, re, im, first=1)
    o = c.beam
    plane = c.ngrid * c.ngrid
    with api.Context(0) as ctx:
        ctx.set_variant(TUNE_ADI, variant)
        f = gpu_field(ctx, c, kind=api.ADI, re=re, im=im, first=1)
        b = gpu_beam(ctx, c)
        for it in range(2):
            f.track(b, st, st.delz)
            sre = np.zeros(c.nslice * plane)
            sim = np.zeros_like(sre)
            on = _oracle.lib().orc_deposit(c.nslice, ip(o.offsets), dp(o.x), dp(o.y),
                                           dp(o.gamma), dp(o.theta), of.ptr, st, dp(o.current),
                                           st.delz, dp(sre), dp(sim))
            assert on == 1
            _oracle.lib().orc_adi_advance(of.ptr, dp(sre), dp(sim), st.delz)
            gre, gim, _ = f.download()
            e = relerr(gre + 1j * gim, of.re + 1j * of.im)
            print(f"  step {it}: rel err {e:.3e}", flush=True)
            assert e < 1e-12
        f.close()
        b.close()


@pytest.mark.parametrize("prec", [api.FP64, api.MIXED])
@pytest.mark.parametrize("ngrid", [45, 151, 256])
def test_segmented_is_the_default(prec, ngrid):
    """Unpinned on G <= 256: the segmented sweeps, equal bit for bit to a
    run pinned to variant 3 and within round-off of the single-chain
    variant 1 (fields and field sums of every record)."""
    print(f"case segmented_default prec={prec} ngrid={ngrid}", flush=True)
    c = cases.make_case(nslice=4, npart=64, ngrid=ngrid, seed_power=1e6, current=0.0, nsteps=6)
    re, im = noisy_field(c, scale=0.1)
    d = _run_records(prec, -1, c, re, im, 6)
    p3 = _run_records(prec, 3, c, re, im, 6)
    p1 = _run_records(prec, 1, c, re, im, 6)
    for x, y in zip(d, p3):
        assert np.array_equal(x, y)
    tol = 1e-13 if prec == api.FP64 else 1e-6
    e = [relerr(x, y) for x, y in zip(d, p1)]
    print(f"  vs single chain: field {e[0]:.2e} {e[1]:.2e} sums {e[2]:.2e}", flush=True)
    assert max(e) < tol


def test_tuner_decides():
    """Unpinned on a grid above the segmented sweeps' limit, the tuner tries
    the variants (three launches each) and then keeps one; the records equal
    a run pinned to variant 0."""
    print("case tuner_decides", flush=True)
    c = cases.make_case(nslice=4, npart=64, ngrid=300, seed_power=1e6, current=0.0, nsteps=12)
    re, im = noisy_field(c, scale=0.1)
    ref = _run_records(api.FP64, 0, c, re, im, 12)
    with api.Context(0) as ctx:
        shape = [c.ngrid, c.nslice, api.FP64]
        assert ctx.get_variant(TUNE_ADI, shape) == -1
        b = gpu_beam(ctx, c)
        f = gpu_field(ctx, c, kind=api.ADI, re=re, im=im, first=2)
        run = api.Run(ctx, b, f, c.steps[:12], LAMBDA0, c.sample, True, False, 1)
        run.advance(6)
        ctx.synchronize()
        run.advance(6)
        _, fraw = run.download_diag()
        gre, gim, _ = f.download()
        v = ctx.get_variant(TUNE_ADI, shape)
        print(f"  decided variant {v}", flush=True)
        assert v in (0, 1, 2)
        run.close()
        f.close()
        b.close()
    assert np.array_equal(gre, ref[0]) and np.array_equal(gim, ref[1])
    assert np.array_equal(fraw, ref[2])


def _split_run(prec, k, c, kind=api.ADI):
    from genesis4_cuda_port_amd._abi import TUNE_SPLIT
    with api.Context(0) as ctx:
        ctx.set_variant(TUNE_SPLIT, k)
        b = gpu_beam(ctx, c, prec)
        f = gpu_field(ctx, c, prec, kind=kind)
        run = api.Run(ctx, b, f, c.steps, LAMBDA0, c.sample, c.timerun, False, 2)
        run.advance(len(c.steps))
        braw, fraw = run.download_diag()
        state = b.download()
        run.close()
        f.close()
        b.close()
    return braw, fraw, state


@pytest.mark.parametrize("kind", [api.ADI, api.FFT])
def test_split_pass_oracle_fp64(kind):
    """A particle pass split over four workgroups per slice (deposition
    windows and sums combined in part order) against the FP64 oracle."""
    print(f"case split_pass_oracle kind={kind}", flush=True)
    from test_gpu_parity import FP64_TOL, _power_bunching
    c = cases.make_case(nslice=3, npart=4096, ngrid=32 if kind == api.FFT else 31, nsteps=20,
                        seed_power=1e4)
    braw, fraw, state = _split_run(api.FP64, 4, c, kind)
    of = OField(c.nslice, c.ngrid, c.gridmax, LAMBDA0, c.slicelength, 1, c.seed_re, c.seed_im)
    ob, obraw, ofraw = _oracle.run(c.beam, of, c.steps, gamma0=cases.GAMMA0, lambda0=LAMBDA0,
                                   solver=kind, sample=c.sample, timerun=c.timerun, nharm=2)
    assert braw.shape == obraw.shape
    # particle counts of every slice are exact
    assert np.array_equal(braw[:, :, 12], obraw[:, :, 12])

    class _F:
        dgrid = 2 * c.gridmax / (c.ngrid - 1)
        xlambda = LAMBDA0
        slicelength = c.slicelength

    P, B = _power_bunching(braw, fraw, c, _F)
    oP, oB = _power_bunching(obraw, ofraw, c, _F)
    ep, eb = relerr(P, oP), relerr(B, oB)
    print(f"  power {ep:.3e} bunching {eb:.3e}", flush=True)
    assert ep < FP64_TOL and eb < FP64_TOL
    g = state[4]
    assert relerr(g - cases.GAMMA0, ob.gamma - cases.GAMMA0) < 1e-10


def test_split_factors_agree_mixed():
    """MIXED: one, three and eight workgroups per slice agree to round-off."""
    print("case split_factors_mixed", flush=True)
    c = cases.make_case(nslice=2, npart=8192, ngrid=33, nsteps=12, seed_power=1e4)
    ref = _split_run(api.MIXED, 1, c)
    for k in (3, 8):
        got = _split_run(api.MIXED, k, c)
        e_b = relerr(got[0], ref[0])
        e_f = relerr(got[1], ref[1])
        print(f"  K={k}: beam raw {e_b:.3e} field raw {e_f:.3e}", flush=True)
        assert e_b < 1e-6 and e_f < 1e-6
        assert np.array_equal(got[0][:, :, 12], ref[0][:, :, 12])


def test_split_automatic_fp64():
    """Unpinned, a two-slice beam of 16384 particles per slice is split; the
    records equal a run pinned to one workgroup per slice to FP64 round-off."""
    print("case split_automatic", flush=True)
    c = cases.make_case(nslice=2, npart=16384, ngrid=33, nsteps=8, seed_power=1e4)
    ref = _split_run(api.FP64, 1, c)
    got = _split_run(api.FP64, -1, c)
    assert relerr(got[0], ref[0]) < 1e-12 and relerr(got[1], ref[1]) < 1e-12
