"""CPU tests of the C ABI surface and the host-only front-end.

No compute call reaches a GPU here: the library is loaded, its exports are
checked against include/*.h, the host-only front-end functions are pinned to
scipy, and error paths are driven to their status codes.
"""
from __future__ import annotations

import ctypes as C
import re
import subprocess
from pathlib import Path

import numpy as np
import pytest
from scipy import special
from scipy.stats import qmc

from genesis4_cuda_port_amd import _abi, api

ROOT = Path(__file__).resolve().parent.parent


def _declared_symbols():
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        text = h.read_text()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names |= set(re.findall(r"\b(g4h_[a-z0-9_]+)\s*\(", text))
    return names


def test_library_exports_every_declared_symbol(g4lib):
    names = _declared_symbols()
    assert len(names) > 30
    out = subprocess.run(["nm", "-D", "--defined-only", str(_abi.LIB_PATH)], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r"\b(g4h_[a-z0-9_]+)\b", out))
    missing = names - exported
    assert not missing, missing
    # and the Python binding knows all of them
    assert not (names - set(_abi.PROTOTYPES)), names - set(_abi.PROTOTYPES)


def test_library_is_gfx950_only():
    """The embedded code objects target gfx950 and nothing else."""
    data = _abi.LIB_PATH.read_bytes()
    targets = set(re.findall(rb"amdgcn-amd-amdhsa--(gfx[0-9a-f]+)", data))
    assert targets == {b"gfx950"}, targets


def test_abi_version_and_no_device_status(g4lib):
    assert g4lib.g4h_abi_version() == 1
    if api.device_count() == 0:
        h = C.c_void_p()
        st = g4lib.g4h_context_create(0, C.byref(h))
        assert st == 6  # G4H_ERR_NO_DEVICE
        assert b"no GPU" in g4lib.g4h_last_error()


def test_null_and_bad_sizes_return_status(g4lib):
    q = np.zeros(100)
    g = np.zeros(10)
    dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))  # noqa: E731
    assert g4lib.g4h_diag_finalize_beam(2, 1, None, None, 11000.0, dp(q), dp(g)) == 1
    assert g4lib.g4h_diag_finalize_field(0, 1e-6, 1e-9, 1e-9, dp(q), dp(q), dp(g)) == 1
    assert g4lib.g4h_sobol(9, 0, 4, dp(q)) == 1
    n = C.c_int()
    aw = np.array([1.0])
    assert g4lib.g4h_lattice_fodo(0.03, dp(aw), 1, 10, 5, 6, 1.0, 1, None, 0, C.byref(n)) == 1
    spec = _abi.BeamSpec()
    spec.nslice, spec.npart, spec.nbins, spec.gamma0 = 1, 10, 4, 100.0  # 10 % 4 != 0
    spec.betax = spec.betay = 1.0
    arrs = [np.zeros(10) for _ in range(6)]
    assert g4lib.g4h_load_beam(C.byref(spec), None, *map(dp, arrs)) == 1
    assert g4lib.g4h_context_create(0, None) == 1
    assert g4lib.g4h_track_transverse(None, None, None, 0.1, 0) == 1
    assert g4lib.g4h_field_track(None, None, None, None, 0.1) == 1
    assert g4lib.g4h_run_advance(None, 1, None) == 1
    assert g4lib.g4h_beam_set_particle_ez(None, None) == 1
    assert g4lib.g4h_beam_add_energy(None, None, None) == 1


@pytest.mark.parametrize("dims", [1, 3, 6, 8])
def test_sobol_matches_scipy(dims):
    n = 1 << 10
    ours = api.sobol(dims, n)
    ref = qmc.Sobol(d=dims, scramble=False).random(n)
    assert np.array_equal(ours, ref)
    # an arbitrary offset index: same as a slice of the sequence
    assert np.array_equal(api.sobol(dims, 100, 37), ref[37:137])


def test_inverse_normal_matches_ndtri():
    p = np.concatenate([np.linspace(1e-12, 1e-3, 50), np.linspace(0.001, 0.999, 999),
                        1 - np.linspace(1e-12, 1e-3, 50)])
    ours = np.array([api.inverse_normal(v) for v in p])
    ref = special.ndtri(p)
    assert np.max(np.abs(ours - ref) / np.maximum(1, np.abs(ref))) < 1e-14


def test_lattice_fodo_table():
    steps = api.lattice_fodo(0.026, [1.7, 1.6], 50, 46, 3, 1.5, 2)
    assert len(steps) == 100
    aw = np.array([s.aw for s in steps])
    assert np.all(aw[:50] == 1.7) and np.all(aw[50:] == 1.6)
    qf = np.array([s.qf for s in steps])
    assert list(np.nonzero(qf)[0]) == [21, 22, 23, 67, 68, 69]
    assert np.all(qf[21:24] == 1.5) and np.all(qf[67:70] == -1.5)
    ku = 2 * np.pi / 0.026
    assert all(s.ku == ku and s.ky == ku * ku and s.kx == 0 for s in steps)
    assert all(s.slippage == 1.0 and s.delz == 0.026 and s.in_undulator == 1 for s in steps)
    assert [s.outstep for s in steps[:4]] == [0, 1, 0, 1]
    assert steps[-1].z == pytest.approx(100 * 0.026)


def test_loader_mirroring_and_thread_independence(monkeypatch):
    kw = dict(nslice=64, npart=1024, nbins=4, gamma0=11350.0, espread_rel=1e-4, emitn=0.4e-6,
              beta=15.0, slicelength=4e-10, current=3000.0, shotnoise=False, seed=5)
    a = api.load_beam(**kw)
    th = a.theta.reshape(-1, 4)
    d = np.diff(th, axis=1)
    assert np.allclose(d, np.pi / 2, atol=1e-15)
    for k in ("x", "px", "gamma"):
        v = getattr(a, k).reshape(-1, 4)
        assert np.all(v == v[:, :1])
    monkeypatch.setenv("OMP_NUM_THREADS", "1")
    b = api.load_beam(**kw)
    monkeypatch.setenv("OMP_NUM_THREADS", "7")
    c = api.load_beam(**kw)
    for k in ("x", "y", "px", "py", "gamma", "theta"):
        assert np.array_equal(getattr(a, k), getattr(b, k))
        assert np.array_equal(getattr(a, k), getattr(c, k))


def test_loader_chirp_and_seed():
    kw = dict(nslice=11, npart=4096, nbins=4, gamma0=10000.0, espread_rel=0.0, emitn=0.4e-6,
              beta=15.0, slicelength=4e-10, current=3000.0, shotnoise=False, seed=5, chirp_rel=1e-3)
    a = api.load_beam(**kw)
    g = a.gamma.reshape(11, -1)
    assert np.allclose(g[:, 0], 10000.0 * (1 + 1e-3 * (np.arange(11) / 10 - 0.5)), rtol=1e-15)
    b = api.load_beam(**dict(kw, seed=6))
    assert not np.array_equal(a.x, b.x)


def test_seed_gaussian_normalisation():
    re, im = api.seed_gaussian(3, 51, 0.3e-3, 4e-10, 1.0e3, 50e-6)
    dg = 0.6e-3 / 50
    ks = 2 * np.pi / 4e-10
    p = np.sum(re[:51 * 51] ** 2 + im[:51 * 51] ** 2) * (dg * 510999.06 / ks) ** 2 / 376.73
    assert p == pytest.approx(1.0e3, rel=1e-12)
    assert np.all(im == 0) and np.array_equal(re[:2601], re[2601:5202])


def test_server_and_batch_need_a_device(g4lib):
    if api.device_count() != 0:
        pytest.skip("a GPU is visible")
    h = C.c_void_p()
    assert g4lib.g4h_server_create(0, 1 << 30, 2, C.byref(h)) == 6  # G4H_ERR_NO_DEVICE
    assert g4lib.g4h_server_create(0, 1 << 30, 0, C.byref(h)) == 1  # bad lane count
    assert g4lib.g4h_batch_create(None, 0, 0, 1, None, None, 1, 1e-9, 1, C.byref(h)) == 1


def test_adapter_layer_compiles_and_links():
    """The reference-shaped C++ layer (csrc/genesis_gpu.hpp) is compiled into
    _lib/g4adapter_check by the build; without a GPU it reports that as an
    error status, not a crash."""
    import subprocess
    from pathlib import Path
    exe = Path(api.__file__).resolve().parent / "_lib" / "g4adapter_check"
    assert exe.exists()
    res = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert res.returncode == 2 and "usage" in res.stderr
    if api.device_count() == 0:
        res = subprocess.run([str(exe), "fp64", "adi", "1", "1"], capture_output=True, text=True,
                             timeout=60)
        assert res.returncode == 1 and "no GPU" in res.stderr
