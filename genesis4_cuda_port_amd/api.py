"""Python handles over the g4hip C ABI (numpy in, numpy out).

These are thin: every method is one ABI call.  The names follow the
reference's classes so a test reads like the reference's call sites:
Beam.track_transverse = TrackBeamCUDA::track, Beam.advance =
BeamSolverCUDA::advance, Field.track = Field::track (solver init + advance),
Field.apply_slippage = ControlCUDA::applySlippage, diag_* =
DiagnosticCUDA::calc.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _abi
from ._abi import (ADI, FFT, FIELD_Q, FIELD_RAW, FP64, MIXED, BeamSpec, G4Error, RunConfig, Step,
                   beam_q_len, beam_raw_len, check, lib)

__all__ = [
    "ADI", "FFT", "FP64", "MIXED", "G4Error", "Step", "Context", "Beam", "Field", "Run",
    "steps_array", "lattice_fodo", "load_beam", "seed_gaussian", "sobol", "undulator_fc",
    "inverse_normal", "finalize_beam", "finalize_field", "device_count",
]


def _dp(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_int32))


def _f64(a, n=None) -> np.ndarray:
    a = np.ascontiguousarray(a, dtype=np.float64)
    if n is not None and a.size != n:
        raise ValueError(f"expected {n} values, got {a.size}")
    return a


def device_count() -> int:
    return int(lib().g4h_device_count())


def steps_array(steps) -> C.Array:
    """A ctypes array of Step from a list of Step (or dicts)."""
    arr = (Step * len(steps))()
    for i, s in enumerate(steps):
        if isinstance(s, dict):
            for k, v in s.items():
                setattr(arr[i], k, v)
        else:
            arr[i] = s
    return arr


class Context:
    def __init__(self, device: int = 0):
        h = C.c_void_p()
        check(lib().g4h_context_create(device, C.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            lib().g4h_context_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def synchronize(self):
        check(lib().g4h_context_synchronize(self.h))

    def memory(self) -> int:
        v = C.c_uint64()
        check(lib().g4h_context_memory(self.h, C.byref(v)))
        return int(v.value)

    def set_variant(self, family: int, variant: int):
        """Pin a kernel family to one variant (-1: autotune)."""
        check(lib().g4h_context_set_variant(self.h, family, variant))

    def get_variant(self, family: int, shape) -> int:
        """Variant the tuner decided for `shape` (-1: still measuring)."""
        arr = (C.c_int * len(shape))(*shape)
        v = C.c_int()
        check(lib().g4h_context_get_variant(self.h, family, arr, len(shape), C.byref(v)))
        return int(v.value)

    def profile(self, on: bool):
        check(lib().g4h_profile_enable(self.h, 1 if on else 0))

    def profile_read(self, cls: int):
        n = C.c_int64()
        ms = C.c_double()
        check(lib().g4h_profile_read(self.h, cls, C.byref(n), C.byref(ms)))
        return int(n.value), float(ms.value)


class Beam:
    def __init__(self, ctx: Context, precision: int, slice_offsets, current, gamma_ref: float):
        off = np.ascontiguousarray(slice_offsets, dtype=np.int32)
        self.nslice = off.size - 1
        self.np = int(off[-1])
        self.off = off
        self.gamma_ref = float(gamma_ref)
        self.precision = precision
        cur = _f64(current, self.nslice)
        self.current = cur
        h = C.c_void_p()
        check(lib().g4h_beam_create(ctx.h, precision, self.nslice, _ip(off), _dp(cur),
                                    self.gamma_ref, C.byref(h)))
        self.h = h
        self.ctx = ctx

    def close(self):
        if self.h:
            lib().g4h_beam_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def upload(self, x, y, px, py, gamma, theta):
        arrs = [_f64(a, self.np) for a in (x, y, px, py, gamma, theta)]
        check(lib().g4h_beam_upload(self.h, *map(_dp, arrs)))

    def download(self):
        arrs = [np.empty(self.np) for _ in range(6)]
        check(lib().g4h_beam_download(self.h, *map(_dp, arrs)))
        return tuple(arrs)

    def generate(self, **kw):
        """Quiet-start beam generated on the device (g4h_beam_generate);
        keywords of beam_spec."""
        spec, cur = beam_spec(**kw)
        check(lib().g4h_beam_generate(self.h, C.byref(spec), _dp(cur)))

    def track_transverse(self, step: Step, delz: float, last: bool):
        check(lib().g4h_track_transverse(self.ctx.h, self.h, C.byref(step), delz, int(last)))

    def apply_r56(self, step: Step, lambda0: float):
        check(lib().g4h_apply_r56(self.ctx.h, self.h, C.byref(step), lambda0))

    def advance(self, fields, step: Step, delz: float, slice_eloss=None):
        arr = (C.c_void_p * len(fields))(*[f.h.value for f in fields])
        el = None if slice_eloss is None else _dp(_f64(slice_eloss, self.nslice))
        check(lib().g4h_beam_advance(self.ctx.h, self.h, arr, len(fields), C.byref(step), delz, el))

    def set_current(self, current):
        """Per-slice current (A), as the source scale and the diagnostics read it."""
        cur = _f64(current, self.nslice)
        check(lib().g4h_beam_set_current(self.h, _dp(cur)))
        self.current = cur

    def info(self):
        """(nslice, total particles, precision) as the library holds them."""
        ns, npart, prec = C.c_int(), C.c_int64(), C.c_int()
        check(lib().g4h_beam_info(self.h, C.byref(ns), C.byref(npart), C.byref(prec)))
        return int(ns.value), int(npart.value), int(prec.value)

    def add_energy(self, dgamma):
        """gamma += dgamma[slice] on the device (nslice values)."""
        check(lib().g4h_beam_add_energy(self.ctx.h, self.h, _dp(_f64(dgamma, self.nslice))))

    def set_particle_ez(self, ez=None):
        """Per-particle longitudinal field of later pushes (beam order, np
        values, units of slice_eloss); None clears it."""
        el = None if ez is None else _dp(_f64(ez, self.np))
        check(lib().g4h_beam_set_particle_ez(self.h, el))

    def cell_index(self, field: "Field") -> np.ndarray:
        out = np.empty(self.np, dtype=np.int32)
        check(lib().g4h_beam_cell_index(self.ctx.h, self.h, field.h, _ip(out)))
        return out

    def diag_raw(self, nharm: int = 1) -> np.ndarray:
        out = np.empty(self.nslice * beam_raw_len(nharm))
        check(lib().g4h_diag_beam_raw(self.ctx.h, self.h, nharm, _dp(out)))
        return out.reshape(self.nslice, beam_raw_len(nharm))

    def init_efield(self, rmax: float, ngrid: int, nz: int, nphi: int, lam: float,
                    longrange: bool = False):
        """Short-range space charge on the device before every push
        (g4h_beam_init_efield); nz = 0 switches it off."""
        from ._abi import EField
        ef = EField(rmax, ngrid, nz, nphi, int(longrange))
        check(lib().g4h_beam_init_efield(self.h, C.byref(ef), lam))

    def sc_field(self) -> np.ndarray:
        """Mean short-range E_z [V/m] per slice from the last push."""
        out = np.empty(self.nslice)
        check(lib().g4h_beam_sc_field(self.h, _dp(out)))
        return out

    def diag_raw_noalloc(self, nharm: int = 1) -> np.ndarray:
        """diag_raw through the entry point that allocates no device memory."""
        out = np.empty(self.nslice * beam_raw_len(nharm))
        check(lib().g4h_diag_beam_raw_noalloc(self.ctx.h, self.h, nharm, _dp(out)))
        return out.reshape(self.nslice, beam_raw_len(nharm))


class Field:
    def __init__(self, ctx: Context, precision: int, nslice: int, ngrid: int, gridmax: float,
                 xlambda0: float, slicelength: float, harm: int = 1):
        h = C.c_void_p()
        check(lib().g4h_field_create(ctx.h, precision, nslice, ngrid, gridmax, xlambda0,
                                     slicelength, harm, C.byref(h)))
        self.h = h
        self.ctx = ctx
        self.nslice, self.ngrid, self.harm = nslice, ngrid, harm
        self.gridmax = gridmax
        self.dgrid = 2 * gridmax / (ngrid - 1)
        self.xlambda = xlambda0 / harm
        self.slicelength = slicelength
        self.precision = precision

    def close(self):
        if self.h:
            lib().g4h_field_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def cells(self) -> int:
        return self.nslice * self.ngrid * self.ngrid

    def upload(self, re, im, first: int = 0):
        check(lib().g4h_field_upload(self.h, _dp(_f64(re, self.cells)), _dp(_f64(im, self.cells)),
                                     first))

    def fill_planes(self, re, im, first: int = 0):
        """Every ring plane <- one plane (g4h_field_fill_planes)."""
        check(lib().g4h_field_fill_planes(self.h, _dp(_f64(re)), _dp(_f64(im)), first))

    def download(self):
        re = np.empty(self.cells)
        im = np.empty(self.cells)
        first = C.c_int()
        check(lib().g4h_field_download(self.h, _dp(re), _dp(im), C.byref(first)))
        return re, im, int(first.value)

    def copy_from(self, src: "Field"):
        check(lib().g4h_field_copy(self.ctx.h, self.h, src.h))

    def ring(self):
        first = C.c_int()
        acc = C.c_double()
        check(lib().g4h_field_get_ring(self.h, C.byref(first), C.byref(acc)))
        return int(first.value), float(acc.value)

    def set_ring(self, first: int, accuslip: float = 0.0):
        check(lib().g4h_field_set_ring(self.h, first, accuslip))

    def set_enabled(self, on: bool):
        check(lib().g4h_field_set_enabled(self.h, int(on)))

    def init_solver(self, kind: int, source_filter=False, xcut=1.0, ycut=1.0, sigmoid=1.0,
                    low_memory=False):
        check(lib().g4h_field_init_solver(self.h, kind, int(source_filter), xcut, ycut, sigmoid,
                                          int(low_memory)))

    def track(self, beam: Beam, step: Step, delz: float):
        check(lib().g4h_field_track(self.ctx.h, self.h, beam.h, C.byref(step), delz))

    def apply_slippage(self, slippage: float, sample: float, timerun=True, periodic=False) -> int:
        n = C.c_int()
        check(lib().g4h_apply_slippage(self.ctx.h, self.h, slippage, sample, int(timerun),
                                       int(periodic), C.byref(n)))
        return int(n.value)

    def diag_raw(self) -> np.ndarray:
        out = np.empty(self.nslice * FIELD_RAW)
        check(lib().g4h_diag_field_raw(self.ctx.h, self.h, _dp(out)))
        return out.reshape(self.nslice, FIELD_RAW)


class Run:
    """The fused Gencore::run loop on device-resident state."""

    def __init__(self, ctx: Context, beam: Beam, field: Field, steps, lambda0: float,
                 sample: float = 1.0, timerun: bool = True, periodic: bool = False, nharm: int = 1,
                 record_diag: bool = True):
        self.steps = steps if isinstance(steps, C.Array) else steps_array(steps)
        cfg = RunConfig(lambda0, sample, int(timerun), int(periodic), nharm, int(record_diag))
        h = C.c_void_p()
        check(lib().g4h_run_create(ctx.h, beam.h, field.h, self.steps, len(self.steps),
                                   C.byref(cfg), C.byref(h)))
        self.h = h
        self.ctx, self.beam, self.field, self.nharm = ctx, beam, field, nharm

    def close(self):
        if self.h:
            lib().g4h_run_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def advance(self, nsteps: int) -> int:
        done = C.c_int()
        check(lib().g4h_run_advance(self.h, nsteps, C.byref(done)))
        return int(done.value)

    def capture(self):
        """Capture the whole run into a hipGraph (after one full advance)."""
        check(lib().g4h_run_capture(self.h))

    def replay(self):
        """Run again from re-initialised beam and field, as one graph launch."""
        check(lib().g4h_run_replay(self.h))

    def records(self) -> int:
        n = C.c_int()
        check(lib().g4h_run_records(self.h, C.byref(n)))
        return int(n.value)

    def download_sc(self) -> np.ndarray:
        """Mean short-range space-charge field per record and slice [V/m]."""
        n = self.records()
        out = np.empty(n * self.beam.nslice)
        check(lib().g4h_run_download_sc(self.h, _dp(out)))
        return out.reshape(n, self.beam.nslice)

    def download_diag(self):
        n = self.records()
        S = self.beam.nslice
        braw = np.empty(n * S * beam_raw_len(self.nharm))
        fraw = np.empty(n * S * FIELD_RAW)
        check(lib().g4h_run_download_diag(self.h, _dp(braw), _dp(fraw)))
        return (braw.reshape(n, S, beam_raw_len(self.nharm)), fraw.reshape(n, S, FIELD_RAW))


class Batch:
    """Independent steady-state runs advanced together (g4h_batch_*).

    runs: list of dicts with keys npart, nbins, ngrid, gridmax, current,
    gamma0, espread_rel, emitn (or emitn_x/emitn_y), beta (or betax/betay),
    alphax, alphay, shotnoise, seed, seed_power, seed_w0 and optionally aw
    (per-step taper, one value per step)."""

    def __init__(self, ctx: Context, precision: int, solver: int, runs, steps, lambda0: float,
                 nharm: int = 1):
        from ._abi import BatchRun
        self.steps = steps if isinstance(steps, C.Array) else steps_array(steps)
        ns = len(self.steps)
        arr = (BatchRun * len(runs))()
        self._aw = []
        for i, r in enumerate(runs):
            q = arr[i]
            q.npart, q.nbins, q.ngrid = r["npart"], r.get("nbins", 4), r["ngrid"]
            q.shotnoise, q.seed = int(r.get("shotnoise", True)), int(r.get("seed", 1))
            q.gridmax, q.current, q.gamma0 = r["gridmax"], r["current"], r["gamma0"]
            q.espread_rel = r.get("espread_rel", 0.0)
            q.emitn_x = r.get("emitn_x", r.get("emitn", 0.4e-6))
            q.emitn_y = r.get("emitn_y", r.get("emitn", 0.4e-6))
            q.betax = r.get("betax", r.get("beta", 15.0))
            q.betay = r.get("betay", r.get("beta", 15.0))
            q.alphax, q.alphay = r.get("alphax", 0.0), r.get("alphay", 0.0)
            q.seed_power, q.seed_w0 = r.get("seed_power", 0.0), r.get("seed_w0", 50e-6)
            if r.get("aw") is not None:
                a = _f64(r["aw"], ns)
                self._aw.append(a)
                q.aw = _dp(a)
        h = C.c_void_p()
        check(lib().g4h_batch_create(ctx.h, precision, solver, len(runs), arr, self.steps, ns,
                                     lambda0, nharm, C.byref(h)))
        self.h, self.ctx, self.nharm, self.nruns = h, ctx, nharm, len(runs)

    def close(self):
        if self.h:
            lib().g4h_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def advance(self, nsteps: int) -> int:
        done = C.c_int()
        check(lib().g4h_batch_advance(self.h, nsteps, C.byref(done)))
        return int(done.value)

    def records(self) -> int:
        n = C.c_int()
        check(lib().g4h_batch_records(self.h, C.byref(n)))
        return int(n.value)

    def download(self, run: int):
        """(beam_raw[records, RAW], field_raw[records, 9]) of one run."""
        n = self.records()
        braw = np.empty(n * beam_raw_len(self.nharm))
        fraw = np.empty(n * FIELD_RAW)
        check(lib().g4h_batch_download(self.h, run, _dp(braw), _dp(fraw)))
        return braw.reshape(n, beam_raw_len(self.nharm)), fraw.reshape(n, FIELD_RAW)


class Server:
    """Request loop keeping many runs resident on one GPU (g4h_server_*)."""

    def __init__(self, device: int = 0, hbm_budget: int = 200 << 30, nstreams: int = 4):
        h = C.c_void_p()
        check(lib().g4h_server_create(device, hbm_budget, nstreams, C.byref(h)))
        self.h = h
        self._meta = {}

    def close(self):
        if self.h:
            lib().g4h_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, *, kind: int, precision: int, solver: int, steps, nslice: int, npart: int,
               ngrid: int, gridmax: float, current: float, gamma0: float, lambda0: float,
               espread_rel: float = 0.0, chirp_rel: float = 0.0, emitn: float = 0.4e-6,
               beta: float = 15.0, nbins: int = 4, shotnoise: bool = True, seed: int = 1,
               seed_power: float = 0.0, seed_w0: float = 50e-6, sample: float = 1.0,
               nharm: int = 1, periodic: bool = False, aw=None) -> int:
        from ._abi import Request
        st = steps if isinstance(steps, C.Array) else steps_array(steps)
        q = Request()
        q.kind, q.precision, q.solver = kind, precision, solver
        q.nslice, q.npart, q.nbins, q.ngrid = nslice, npart, nbins, ngrid
        q.shotnoise, q.nharm, q.periodic, q.nsteps = int(shotnoise), nharm, int(periodic), len(st)
        q.seed, q.gridmax, q.current, q.gamma0 = seed, gridmax, current, gamma0
        q.espread_rel, q.chirp_rel, q.emitn, q.beta = espread_rel, chirp_rel, emitn, beta
        q.seed_power, q.seed_w0, q.lambda0, q.sample = seed_power, seed_w0, lambda0, sample
        q.steps = st
        a = None
        if aw is not None:
            a = _f64(aw, len(st))
            q.aw = _dp(a)
        i = C.c_int64()
        check(lib().g4h_server_submit(self.h, C.byref(q), C.byref(i)))
        self._meta[int(i.value)] = (nslice, nharm)
        return int(i.value)

    def wait(self, rid: int, timeout: float = -1.0):
        from ._abi import ResultInfo
        info = ResultInfo()
        check(lib().g4h_server_wait(self.h, rid, timeout, C.byref(info)))
        return info

    def result(self, rid: int):
        info = self.wait(rid)
        S, nh = self._meta[rid]
        n = info.nrecords
        braw = np.empty(n * S * beam_raw_len(nh))
        fraw = np.empty(n * S * FIELD_RAW)
        check(lib().g4h_server_result(self.h, rid, _dp(braw), _dp(fraw)))
        return braw.reshape(n, S, beam_raw_len(nh)), fraw.reshape(n, S, FIELD_RAW)

    def release(self, rid: int):
        check(lib().g4h_server_release(self.h, rid))
        self._meta.pop(rid, None)

    def stats(self):
        from ._abi import ServerStats
        st = ServerStats()
        check(lib().g4h_server_stats(self.h, C.byref(st)))
        return {"completed": st.completed, "p50_s": st.p50_s, "p99_s": st.p99_s,
                "mean_s": st.mean_s, "particle_slice_steps_per_s": st.particle_slice_steps_per_s,
                "hbm_in_use": st.hbm_in_use}


# ---- front-end --------------------------------------------------------------

def undulator_fc(aw: float, helical: bool, harm: int) -> float:
    return float(lib().g4h_undulator_fc(aw, int(helical), harm))


def inverse_normal(p: float) -> float:
    return float(lib().g4h_inverse_normal(p))


def sobol(dims: int, n: int, index0: int = 0) -> np.ndarray:
    out = np.empty(n * dims)
    check(lib().g4h_sobol(dims, index0, n, _dp(out)))
    return out.reshape(n, dims)


def lattice_fodo(lambdau: float, aw_mod, nper: int, quad_spacing: int, quad_len: int, k1: float,
                 output_step: int = 1):
    aw = _f64(aw_mod)
    n = C.c_int()
    check(lib().g4h_lattice_fodo(lambdau, _dp(aw), aw.size, nper, quad_spacing, quad_len, k1,
                                 output_step, None, 0, C.byref(n)))
    arr = (Step * n.value)()
    check(lib().g4h_lattice_fodo(lambdau, _dp(aw), aw.size, nper, quad_spacing, quad_len, k1,
                                 output_step, arr, n.value, C.byref(n)))
    return arr


def lattice_parse(text: str, beamline: str, delz: float, lambda0: float, gamma0: float,
                  output_step: int = 1, zstop: float = 1e9):
    """Per-step table of `beamline` in lattice file text (g4h_lattice_parse)."""
    n = C.c_int()
    args = (text.encode(), beamline.encode(), delz, lambda0, gamma0, output_step, zstop)
    check(lib().g4h_lattice_parse(*args, None, 0, C.byref(n)))
    arr = (Step * n.value)()
    check(lib().g4h_lattice_parse(*args, arr, n.value, C.byref(n)))
    return arr


def lattice_match(steps, gamma0: float, zmatch: float):
    """Matched (betax, alphax, betay, alphay) of the section [0, zmatch]."""
    t = np.empty(4)
    check(lib().g4h_lattice_match(steps, len(steps), gamma0, zmatch, _dp(t)))
    return tuple(float(v) for v in t)


def deck_parse(text: str):
    """Namelists of an input deck, in file order: [(name, [(key, value), ...]), ...]."""
    import json
    need = C.c_int()
    check(lib().g4h_deck_parse(text.encode(), None, 0, C.byref(need)))
    buf = C.create_string_buffer(need.value)
    check(lib().g4h_deck_parse(text.encode(), buf, need.value, C.byref(need)))
    return [(d["name"], [tuple(kv) for kv in d["keys"]]) for d in json.loads(buf.value.decode())]


def deck_run(deck_path: str, out_dir: str, device: int = 0) -> int:
    """Run an input deck on the GPU; returns the number of records written."""
    n = C.c_int()
    check(lib().g4h_deck_run(str(deck_path).encode(), str(out_dir).encode(), device,
                             C.byref(n)))
    return int(n.value)


@dataclass
class BeamArrays:
    x: np.ndarray
    y: np.ndarray
    px: np.ndarray
    py: np.ndarray
    gamma: np.ndarray
    theta: np.ndarray
    offsets: np.ndarray
    current: np.ndarray

    def copy(self) -> "BeamArrays":
        return BeamArrays(*(getattr(self, k).copy() for k in
                            ("x", "y", "px", "py", "gamma", "theta", "offsets", "current")))


def beam_spec(*, nslice: int, npart: int, nbins: int, gamma0: float, espread_rel: float,
              emitn: float, beta: float, alpha: float = 0.0, slicelength: float, current,
              shotnoise: bool, seed: int, chirp_rel: float = 0.0, emitn_y: float | None = None,
              beta_y: float | None = None, alpha_y: float | None = None):
    """(g4h_beam_spec, per-slice currents) of a quiet-start beam; the y plane
    defaults to the x plane."""
    cur = _f64(np.broadcast_to(np.asarray(current, dtype=np.float64), (nslice,)).copy(), nslice)
    ey = emitn if emitn_y is None else emitn_y
    by = beta if beta_y is None else beta_y
    ay = alpha if alpha_y is None else alpha_y
    spec = BeamSpec(gamma0, espread_rel, chirp_rel, emitn, ey, beta, by, alpha, ay,
                    0.0, 0.0, 0.0, 0.0, slicelength, nslice, npart, nbins, int(shotnoise), seed)
    return spec, cur


def load_beam(**kw) -> BeamArrays:
    """Quiet-start beam on the host (g4h_load_beam); keywords of beam_spec."""
    spec, cur = beam_spec(**kw)
    nslice, npart = spec.nslice, spec.npart
    n = nslice * npart
    arrs = [np.empty(n) for _ in range(6)]
    check(lib().g4h_load_beam(C.byref(spec), _dp(cur), *map(_dp, arrs)))
    off = (np.arange(nslice + 1, dtype=np.int64) * npart).astype(np.int32)
    return BeamArrays(*arrs, off, cur)


def seed_gaussian(nslice: int, ngrid: int, gridmax: float, xlambda: float, power: float, w0: float):
    cells = nslice * ngrid * ngrid
    re = np.empty(cells)
    im = np.empty(cells)
    check(lib().g4h_seed_gaussian(nslice, ngrid, gridmax, xlambda, power, w0, _dp(re), _dp(im)))
    return re, im


def finalize_beam(raw: np.ndarray, nharm: int, current, gamma_ref: float):
    raw = _f64(raw)
    S = raw.size // beam_raw_len(nharm)
    q = np.empty(S * beam_q_len(nharm))
    g = np.empty(6)
    check(lib().g4h_diag_finalize_beam(S, nharm, _dp(raw), _dp(_f64(current, S)), gamma_ref,
                                       _dp(q), _dp(g)))
    return q.reshape(S, beam_q_len(nharm)), g


def finalize_field(raw: np.ndarray, dgrid: float, xlambda: float, slicelength: float):
    raw = _f64(raw)
    S = raw.size // FIELD_RAW
    q = np.empty(S * FIELD_Q)
    g = np.empty(7)
    check(lib().g4h_diag_finalize_field(S, dgrid, xlambda, slicelength, _dp(raw), _dp(q), _dp(g)))
    return q.reshape(S, FIELD_Q), g
