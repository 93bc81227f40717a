"""GPU: the reference's example deck (test/fel.in + test/fel.lat, shortened
time window and lattice) run end to end through g4h_deck_run, against the
FP64 oracle fed with the same step table, beam and field.

Bar: FP64 slice power and bunching within 1e-10 relative, as for the other
FP64 runs."""
from __future__ import annotations

import json
import shutil
from pathlib import Path

import numpy as np
import pytest

import _oracle
from _oracle import OField
from genesis4_cuda_port_amd import api

pytestmark = pytest.mark.gpu

EX = Path(__file__).resolve().parent / "golden" / "reference_example"
LAMBDA0, GAMMA0, SAMPLE, ZSTOP = 0.15e-9, 27984.34, 8, 6.0
NSLICE = 10


def relerr(a, b):
    den = np.linalg.norm(np.ravel(b))
    return np.linalg.norm(np.ravel(a) - np.ravel(b)) / (den if den > 0 else 1.0)


def _deck(tmp_path: Path, solver_keys: str) -> Path:
    text = (EX / "fel.in").read_text()
    text = text.replace("slen=2.0e-6", f"slen={NSLICE * SAMPLE * LAMBDA0!r}")
    text = text.replace("use_cuda=1", f"use_cuda=1\nzstop={ZSTOP}\nprecision=fp64\n{solver_keys}")
    shutil.copy(EX / "fel.lat", tmp_path / "fel.lat")
    deck = tmp_path / "fel.in"
    deck.write_text(text)
    return deck


@pytest.mark.parametrize("fft", [False, True])
def test_reference_deck_matches_oracle(tmp_path, fft):
    print(f"case deck fft={fft}", flush=True)
    deck = _deck(tmp_path, "fft_fieldsolver=1" if fft else "")
    out = tmp_path / "out"
    nrec = api.deck_run(deck, out)
    meta = json.loads((out / "meta.json").read_text())
    assert meta["nslice"] == NSLICE and meta["nrecords"] == nrec
    # the same inputs, built through the front-end calls and run on the oracle
    steps = api.lattice_parse((EX / "fel.lat").read_text(), "FEL", 0.24, LAMBDA0, GAMMA0, 1, ZSTOP)
    assert nrec == 1 + len(steps)
    beam = api.load_beam(nslice=NSLICE, npart=4800, nbins=4, gamma0=GAMMA0,
                         espread_rel=2.24 / GAMMA0, emitn=1.2e-6, beta=meta["betax"],
                         alpha=meta["alphax"], beta_y=meta["betay"], alpha_y=meta["alphay"],
                         slicelength=SAMPLE * LAMBDA0, current=3400.0, shotnoise=True,
                         seed=1234567)
    bx, ax, by, ay = api.lattice_match(api.lattice_parse((EX / "fel.lat").read_text(), "FEL",
                                                         0.24, LAMBDA0, GAMMA0), GAMMA0, 5.72)
    assert (bx, ax, by, ay) == pytest.approx((meta["betax"], meta["alphax"], meta["betay"],
                                              meta["alphay"]), rel=1e-12, abs=1e-12)
    of = OField(NSLICE, 127, 120e-6, LAMBDA0, SAMPLE * LAMBDA0)
    _, braw, fraw = _oracle.run(beam, of, steps, gamma0=GAMMA0, lambda0=LAMBDA0,
                                solver=api.FFT if fft else api.ADI, sample=SAMPLE, timerun=True,
                                nharm=1)
    dgrid = 2 * 120e-6 / 126
    P = np.array([api.finalize_field(fraw[r], dgrid, LAMBDA0, SAMPLE * LAMBDA0)[0][:, 0]
                  for r in range(nrec)])
    B = np.array([api.finalize_beam(braw[r], 1, np.full(NSLICE, 3400.0), GAMMA0)[0][:, 25]
                  for r in range(nrec)])
    gp = np.load(out / "Field" / "power.npy")
    gb = np.load(out / "Beam" / "bunching.npy")
    assert gp.shape == (nrec, NSLICE) and gb.shape == (nrec, NSLICE)
    ep, eb = relerr(gp[1:], P[1:]), relerr(gb, B)
    print(f"  power rel err {ep:.3e}  bunching rel err {eb:.3e}", flush=True)
    # the tail slice takes the empty incoming plane at every slippage shift
    assert gp[-1, 0] == 0 and np.all(gp[-1, 1:] > 0)
    assert ep < 1e-10 and eb < 1e-10
    z = np.load(out / "Global_zplot.npy").ravel()
    assert z[0] == 0 and z[-1] == pytest.approx(steps[len(steps) - 1].z)


def test_g4run_cli_writes_the_same_records(tmp_path):
    import subprocess
    print("case deck cli", flush=True)
    deck = _deck(tmp_path, "")
    cli = Path(api.__file__).resolve().parent / "_lib" / "g4run"
    res = subprocess.run([str(cli), str(deck), "-o", str(tmp_path / "cli")], capture_output=True,
                         text=True, timeout=300)
    assert res.returncode == 0, res.stderr
    nrec = api.deck_run(deck, tmp_path / "lib")
    assert f"{nrec} records" in res.stdout
    for name in ("Field/power", "Beam/bunching", "Beam/energy", "Field/xsize"):
        a = np.load(tmp_path / "cli" / f"{name}.npy")
        b = np.load(tmp_path / "lib" / f"{name}.npy")
        assert relerr(a, b) < 1e-12, name


def test_deck_dumps_match_records(tmp_path):
    """&track field_dump_step / beam_dump_step write the state after those
    integration steps; a field dump gives the slice power of the record of
    the same step, a beam dump the slice energy and particle count."""
    print("case deck dumps", flush=True)
    deck = _deck(tmp_path, "field_dump_step=10\nbeam_dump_step=15")
    out = tmp_path / "out"
    nrec = api.deck_run(deck, out)
    dumps = sorted(int(p.name) for p in (out / "Dump").iterdir())
    assert dumps == [k for k in range(1, nrec) if k % 10 == 0 or k % 15 == 0]
    gp = np.load(out / "Field" / "power.npy")
    ge = np.load(out / "Beam" / "energy.npy")
    dgrid = 2 * 120e-6 / 126
    ks = 2 * np.pi / LAMBDA0
    for k in dumps:
        meta = json.loads((out / "Dump" / str(k) / "dump.json").read_text())
        assert meta["step"] == k and meta["field"] == (k % 10 == 0)
        if meta["field"]:
            re = np.load(out / "Dump" / str(k) / "field_re.npy")
            im = np.load(out / "Dump" / str(k) / "field_im.npy")
            assert re.shape == (NSLICE, 127 * 127)
            p = np.sum(re * re + im * im, axis=1) * (dgrid * 510999.06 / ks) ** 2 / 376.73
            assert relerr(p, gp[k]) < 1e-10
        if meta["beam"]:
            g = np.load(out / "Dump" / str(k) / "beam_gamma.npy")
            assert g.shape == (NSLICE, 4800)
            assert relerr(g.mean(axis=1), ge[k]) < 1e-12


def test_exclusive_harmonics_writes_fundamental_and_highest(tmp_path):
    """&track exclusive_harmonics = true with bunchharm = 3 writes the
    bunching of harmonics 1 and 3 only (DiagnosticCUDA.cpp:495-507), with
    the values of the run without the switch."""
    print("case deck exclusive_harmonics", flush=True)
    all_ = tmp_path / "all"
    excl = tmp_path / "excl"
    all_.mkdir()
    excl.mkdir()
    nrec = api.deck_run(_deck(all_, "bunchharm=3"), all_ / "out")
    assert api.deck_run(_deck(excl, "bunchharm=3\nexclusive_harmonics=true"), excl / "out") == nrec
    for name in ("bunching", "bunchingphase"):
        for h in ("", "3"):
            a = np.load(excl / "out" / "Beam" / f"{name}{h}.npy")
            b = np.load(all_ / "out" / "Beam" / f"{name}{h}.npy")
            assert a.shape == (nrec, NSLICE) and relerr(a, b) < 1e-12, name + h
        assert (all_ / "out" / "Beam" / f"{name}2.npy").exists()
        assert not (excl / "out" / "Beam" / f"{name}2.npy").exists()


def test_efield_namelist_runs_device_space_charge(tmp_path):
    """&efield (rmax, ngrid, nz, nphi) before &track: the deck's run equals the
    oracle's run with the same short-range space charge, and the records hold
    sscfield (non-zero after the first step) beside zero efield, wakefield and
    lscfield (no wakes or long-range field in this runner)."""
    print("case deck efield", flush=True)
    deck = _deck(tmp_path, "")
    text = deck.read_text().replace("&track", "&efield\nrmax=2.0\nngrid=40\nnz=2\nnphi=1\n&end\n\n"
                                               "&track")
    deck.write_text(text)
    out = tmp_path / "out"
    nrec = api.deck_run(deck, out)
    meta = json.loads((out / "meta.json").read_text())
    steps = api.lattice_parse((EX / "fel.lat").read_text(), "FEL", 0.24, LAMBDA0, GAMMA0, 1, ZSTOP)
    beam = api.load_beam(nslice=NSLICE, npart=4800, nbins=4, gamma0=GAMMA0,
                         espread_rel=2.24 / GAMMA0, emitn=1.2e-6, beta=meta["betax"],
                         alpha=meta["alphax"], beta_y=meta["betay"], alpha_y=meta["alphay"],
                         slicelength=SAMPLE * LAMBDA0, current=3400.0, shotnoise=True,
                         seed=1234567)
    of = OField(NSLICE, 127, 120e-6, LAMBDA0, SAMPLE * LAMBDA0)
    _, braw, fraw = _oracle.run(beam, of, steps, gamma0=GAMMA0, lambda0=LAMBDA0, solver=api.ADI,
                                sample=SAMPLE, timerun=True, nharm=1, efield=(2.0, 40, 2, 1))
    dgrid = 2 * 120e-6 / 126
    P = np.array([api.finalize_field(fraw[r], dgrid, LAMBDA0, SAMPLE * LAMBDA0)[0][:, 0]
                  for r in range(nrec)])
    B = np.array([api.finalize_beam(braw[r], 1, np.full(NSLICE, 3400.0), GAMMA0)[0][:, 25]
                  for r in range(nrec)])
    gp = np.load(out / "Field" / "power.npy")
    gb = np.load(out / "Beam" / "bunching.npy")
    ep, eb = relerr(gp[1:], P[1:]), relerr(gb, B)
    print(f"  power rel err {ep:.3e}  bunching rel err {eb:.3e}", flush=True)
    assert ep < 1e-10 and eb < 1e-10
    ssc = np.load(out / "Beam" / "sscfield.npy")
    assert ssc.shape == (nrec, NSLICE) and np.all(ssc[0] == 0) and np.any(ssc[1:] != 0)
    for name in ("efield", "wakefield", "lscfield"):
        assert np.array_equal(np.load(out / "Beam" / f"{name}.npy"), np.zeros((nrec, NSLICE)))
