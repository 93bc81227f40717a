"""GPU: the C++ adapter layer (csrc/genesis_gpu.hpp, the classes with the
reference's names that GENESIS call sites bind) driven end to end by
_lib/g4adapter_check: its Gencore::run, operator by operator and through
the fused loop, against the FP64 oracle on the same seeded case."""
from __future__ import annotations

import subprocess
from pathlib import Path

import numpy as np
import pytest

import _oracle
from cases import GAMMA0, LAMBDA0, LAMBDAU, AW
from genesis4_cuda_port_amd import api

pytestmark = pytest.mark.gpu

from test_gpu_parity import relerr  # noqa: E402

CHECK = Path(api.__file__).resolve().parent / "_lib" / "g4adapter_check"
NSLICE, NPART, GRIDMAX = 6, 512, 0.3e-3


EF = (2.0, 40, 2, 1)  # the short-range space charge of `g4adapter_check ... sc`


def _adapter(prec, solver, nfield, fused, sc=False):
    print(f"case adapter {prec} {solver} nfield={nfield} fused={fused} sc={sc}", flush=True)
    res = subprocess.run([str(CHECK), prec, solver, str(nfield), str(int(fused))] +
                         (["sc"] if sc else []), capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stderr
    rows = np.array([[float(v) for v in line.split()] for line in res.stdout.splitlines()])
    if sc and not fused:
        return (rows[:, 1:1 + NSLICE], rows[:, 1 + NSLICE:1 + 2 * NSLICE],
                rows[:, 1 + 2 * NSLICE:])
    return rows[:, 1:1 + NSLICE], rows[:, 1 + NSLICE:]


def _case_inputs(ngrid):
    assert abs(LAMBDA0 - LAMBDAU * (1 + AW * AW) / (2 * GAMMA0 ** 2)) < 1e-20
    beam = api.load_beam(nslice=NSLICE, npart=NPART, nbins=4, gamma0=GAMMA0, espread_rel=1e-4,
                         emitn=0.4e-6, beta=15.0, slicelength=LAMBDA0, current=3000.0,
                         shotnoise=True, seed=4242)
    steps = api.lattice_fodo(LAMBDAU, [AW], 12, 46, 3, 1.5, 1)
    re, im = api.seed_gaussian(NSLICE, ngrid, GRIDMAX, LAMBDA0, 1e5, 50e-6)
    return beam, steps, re, im


def _oracle_case(solver, efield=None):
    """The case of g4adapter_check through the oracle."""
    ngrid = 32 if solver == "fft" else 33
    beam, steps, re, im = _case_inputs(ngrid)
    of = _oracle.OField(NSLICE, ngrid, GRIDMAX, LAMBDA0, LAMBDA0, 1, re, im)
    _, braw, fraw = _oracle.run(beam, of, steps, gamma0=GAMMA0, lambda0=LAMBDA0,
                                solver=1 if solver == "fft" else 0, sample=1.0, timerun=True,
                                efield=efield)
    dgrid = 2 * GRIDMAX / (ngrid - 1)
    P, B = [], []
    for k in range(braw.shape[0]):
        fq, _ = api.finalize_field(fraw[k].ravel(), dgrid, LAMBDA0, LAMBDA0)
        bq, _ = api.finalize_beam(braw[k].ravel(), 1, beam.current, GAMMA0)
        P.append(fq[:, 0])
        B.append(bq[:, 25])
    return np.array(P), np.array(B)


@pytest.mark.parametrize("solver", ["adi", "fft"])
def test_adapter_matches_oracle_fp64(solver):
    P, B = _oracle_case(solver)
    for fused in (True, False):
        p, b = _adapter("fp64", solver, 1, fused)
        assert p.shape == P.shape
        ep, eb = relerr(p, P), relerr(b, B)
        print(f"  fused={fused}: power {ep:.2e} bunching {eb:.2e}", flush=True)
        assert ep < 1e-10 and eb < 1e-10


def test_adapter_mixed_and_two_fields():
    P, B = _oracle_case("fft")
    p, b = _adapter("mixed", "fft", 1, False)
    ep, eb = relerr(p, P), relerr(b, B)
    print(f"  mixed: power {ep:.2e} bunching {eb:.2e}", flush=True)
    assert ep < 1e-3 and eb < 1e-3
    # fundamental + third harmonic, operator by operator, against the oracle's
    # operators in the same order (Beam::track, then Field::track and the
    # slippage of each field)
    p2, b2 = _adapter("fp64", "adi", 2, False)
    P2, B2 = _oracle_two_fields()
    ep, eb = relerr(p2, P2), relerr(b2, B2)
    print(f"  two fields: power {ep:.2e} bunching {eb:.2e}", flush=True)
    assert ep < 1e-10 and eb < 1e-10


def _oracle_two_fields():
    """g4adapter_check's two-field case (fundamental seeded, third harmonic
    empty, ADI) through the oracle's operators: per step the transverse half
    step, the push in both fields, R56, the other half step, then for each
    field its source and ADI step, then each field's slippage; the first
    field's power and the bunching recorded every step."""
    from _oracle import OField, OrcField, dp, ip
    beam, steps, re, im = _case_inputs(33)
    b = beam.copy()
    lib = _oracle.lib()
    fields = [OField(NSLICE, 33, GRIDMAX, LAMBDA0, LAMBDA0, 1, re, im),
              OField(NSLICE, 33, GRIDMAX, LAMBDA0, LAMBDA0, 3)]
    np_ = b.x.size
    dgrid = 2 * GRIDMAX / 32
    P, B = [], []

    def record():
        fr = np.zeros(NSLICE * 9)
        lib.orc_diag_field_raw(fields[0].ptr, dp(fr))
        br = np.zeros(NSLICE * 25)
        lib.orc_diag_beam_raw(NSLICE, ip(b.offsets), dp(b.x), dp(b.y), dp(b.px), dp(b.py),
                              dp(b.gamma), dp(b.theta), 1, GAMMA0, dp(br))
        P.append(api.finalize_field(fr, dgrid, LAMBDA0, LAMBDA0)[0][:, 0])
        B.append(api.finalize_beam(br, 1, beam.current, GAMMA0)[0][:, 25])

    record()
    import ctypes as C
    for k in range(len(steps)):
        st = C.pointer(steps[k])
        dz = steps[k].delz
        lib.orc_track_transverse(np_, dp(b.x), dp(b.y), dp(b.px), dp(b.py), dp(b.gamma), st,
                                 GAMMA0, 0.5 * dz, 0)
        arr = (OrcField * 2)(fields[0].s, fields[1].s)
        for i, f in enumerate(fields):
            arr[i].re = dp(f.re)
            arr[i].im = dp(f.im)
        lib.orc_beam_advance(NSLICE, ip(b.offsets), dp(b.x), dp(b.y), dp(b.px), dp(b.py),
                             dp(b.gamma), dp(b.theta), arr, 2, st, GAMMA0, dz, None)
        lib.orc_apply_r56(np_, dp(b.gamma), dp(b.theta), st, GAMMA0, LAMBDA0)
        lib.orc_track_transverse(np_, dp(b.x), dp(b.y), dp(b.px), dp(b.py), dp(b.gamma), st,
                                 GAMMA0, 0.5 * dz, 1)
        for f in fields:
            sre = np.zeros(NSLICE * 33 * 33)
            sim = np.zeros_like(sre)
            on = lib.orc_deposit(NSLICE, ip(b.offsets), dp(b.x), dp(b.y), dp(b.gamma),
                                 dp(b.theta), f.ptr, st, dp(beam.current), dz, dp(sre), dp(sim))
            lib.orc_adi_advance(f.ptr, dp(sre) if on else None, dp(sim) if on else None, dz)
        for f in fields:
            lib.orc_apply_slippage(f.ptr, steps[k].slippage, 1.0, 1, 0)
        if steps[k].outstep:
            record()
    return np.array(P), np.array(B)


def test_adapter_space_charge_matches_oracle():
    """Beam::initEField through the adapter: the fused loop and the
    operator-by-operator loop (BeamSolverGPU::advance) against the oracle's
    run with the same short-range space charge; getSCField per record."""
    P, B = _oracle_case("adi", efield=EF)
    P0, B0 = _oracle_case("adi")
    assert relerr(B, B0) > 1e-6  # space charge changes the run
    p, b = _adapter("fp64", "adi", 1, True, sc=True)
    ep, eb = relerr(p, P), relerr(b, B)
    print(f"  fused: power {ep:.2e} bunching {eb:.2e}", flush=True)
    assert ep < 1e-10 and eb < 1e-10
    p, b, scf = _adapter("fp64", "adi", 1, False, sc=True)
    ep, eb = relerr(p, P), relerr(b, B)
    print(f"  operators: power {ep:.2e} bunching {eb:.2e}", flush=True)
    assert ep < 1e-10 and eb < 1e-10
    assert np.all(scf[0] == 0) and np.all(np.isfinite(scf)) and np.any(scf[1:] != 0)
