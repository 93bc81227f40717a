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


def _adapter(prec, solver, nfield, fused):
    print(f"case adapter {prec} {solver} nfield={nfield} fused={fused}", flush=True)
    res = subprocess.run([str(CHECK), prec, solver, str(nfield), str(int(fused))],
                         capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stderr
    rows = np.array([[float(v) for v in line.split()] for line in res.stdout.splitlines()])
    return rows[:, 1:1 + NSLICE], rows[:, 1 + NSLICE:]


def _oracle_case(solver):
    """The case of g4adapter_check through the oracle."""
    assert abs(LAMBDA0 - LAMBDAU * (1 + AW * AW) / (2 * GAMMA0 ** 2)) < 1e-20
    ngrid = 32 if solver == "fft" else 33
    beam = api.load_beam(nslice=NSLICE, npart=NPART, nbins=4, gamma0=GAMMA0, espread_rel=1e-4,
                         emitn=0.4e-6, beta=15.0, slicelength=LAMBDA0, current=3000.0,
                         shotnoise=True, seed=4242)
    steps = api.lattice_fodo(LAMBDAU, [AW], 12, 46, 3, 1.5, 1)
    re, im = api.seed_gaussian(NSLICE, ngrid, GRIDMAX, LAMBDA0, 1e5, 50e-6)
    of = _oracle.OField(NSLICE, ngrid, GRIDMAX, LAMBDA0, LAMBDA0, 1, re, im)
    _, braw, fraw = _oracle.run(beam, of, steps, gamma0=GAMMA0, lambda0=LAMBDA0,
                                solver=1 if solver == "fft" else 0, sample=1.0, timerun=True)
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
    # fundamental + third harmonic, operator by operator
    p2, b2 = _adapter("fp64", "adi", 2, False)
    assert np.all(np.isfinite(p2)) and np.all(np.isfinite(b2)) and p2[-1].max() > 0
