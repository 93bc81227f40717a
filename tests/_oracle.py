"""ctypes binding of the FP64 CPU oracle (oracle/g4oracle.h).

Test infrastructure: only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg load this library, always as the checker.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

from genesis4_cuda_port_amd._abi import EField, Step, beam_raw_len, FIELD_RAW

ORACLE_LIB = Path(__file__).resolve().parent.parent / "oracle" / "_build" / "libg4oracle.so"


class OrcField(C.Structure):
    _fields_ = [
        ("nslice", C.c_int32), ("ngrid", C.c_int32), ("first", C.c_int32), ("harm", C.c_int32),
        ("enabled", C.c_int32), ("pad_", C.c_int32),
        ("gridmax", C.c_double), ("dgrid", C.c_double), ("xks", C.c_double),
        ("xlambda", C.c_double), ("slicelength", C.c_double), ("accuslip", C.c_double),
        ("re", C.POINTER(C.c_double)), ("im", C.POINTER(C.c_double)),
    ]


_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)
_sp = C.POINTER(Step)
_fp = C.POINTER(OrcField)

PROTO = {
    "orc_version": (C.c_int, []),
    "orc_set_threads": (None, [C.c_int]),
    "orc_fc": (C.c_double, [C.c_double, C.c_int, C.c_int]),
    "orc_track_transverse": (None, [C.c_int64, _dp, _dp, _dp, _dp, _dp, _sp, C.c_double,
                                    C.c_double, C.c_int]),
    "orc_apply_r56": (None, [C.c_int64, _dp, _dp, _sp, C.c_double, C.c_double]),
    "orc_beam_advance": (None, [C.c_int, _ip, _dp, _dp, _dp, _dp, _dp, _dp, _fp, C.c_int, _sp,
                                C.c_double, C.c_double, _dp]),
    "orc_beam_advance_ez": (None, [C.c_int, _ip, _dp, _dp, _dp, _dp, _dp, _dp, _fp, C.c_int,
                                   _sp, C.c_double, C.c_double, _dp]),
    "orc_deposit": (C.c_int, [C.c_int, _ip, _dp, _dp, _dp, _dp, _fp, _sp, _dp, C.c_double, _dp,
                              _dp]),
    "orc_cell_index": (None, [C.c_int64, _dp, _dp, C.c_double, C.c_double, C.c_int, _ip]),
    "orc_adi_coefficients": (None, [C.c_int, C.c_double, _dp, _dp, _dp, _dp, _dp, _dp]),
    "orc_adi_advance": (None, [_fp, _dp, _dp, C.c_double]),
    "orc_fft_advance": (None, [_fp, _dp, _dp, C.c_double, C.c_int, C.c_double, C.c_double,
                               C.c_double]),
    "orc_dft": (None, [C.c_int, _dp, _dp, _dp, _dp, C.c_int]),
    "orc_apply_slippage": (C.c_int, [_fp, C.c_double, C.c_double, C.c_int, C.c_int]),
    "orc_diag_beam_raw": (None, [C.c_int, _ip, _dp, _dp, _dp, _dp, _dp, _dp, C.c_int, C.c_double,
                                 _dp]),
    "orc_diag_field_raw": (None, [_fp, _dp]),
    "orc_finalize_beam": (None, [C.c_int, C.c_int, _dp, _dp, C.c_double, _dp, _dp]),
    "orc_finalize_field": (None, [C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, _dp, _dp,
                                  _dp]),
    "orc_run": (C.c_int, [C.c_int, _ip, _dp, _dp, _dp, _dp, _dp, _dp, C.c_double, C.c_double,
                          _dp, _dp, _fp, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                          _sp, C.c_int, C.c_double, C.c_int, C.c_int, C.c_int, _dp, _dp]),
    "orc_run_sc": (C.c_int, [C.c_int, _ip, _dp, _dp, _dp, _dp, _dp, _dp, C.c_double, C.c_double,
                             _dp, _dp, _fp, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                             _sp, C.c_int, C.c_double, C.c_int, C.c_int, C.c_int, _dp, _dp,
                             C.POINTER(EField)]),
    "orc_sc_field": (None, [C.c_int, _ip, _dp, _dp, _dp, _dp, C.c_double, C.c_double,
                            C.POINTER(EField), _dp, _dp]),
}

_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        if not ORACLE_LIB.exists():
            raise ImportError(f"{ORACLE_LIB} not built (make -C oracle)")
        L = C.CDLL(str(ORACLE_LIB))
        for k, (r, a) in PROTO.items():
            f = getattr(L, k)
            f.restype = r
            f.argtypes = a
        _LIB = L
    return _LIB


def dp(a):
    return a.ctypes.data_as(_dp)


def ip(a):
    return a.ctypes.data_as(_ip)


class OField:
    """Oracle field: storage-order planes in numpy arrays plus ring state."""

    def __init__(self, nslice, ngrid, gridmax, xlambda0, slicelength, harm=1, re=None, im=None,
                 first=0):
        self.re = np.zeros(nslice * ngrid * ngrid) if re is None else np.array(re, dtype=np.float64)
        self.im = np.zeros(nslice * ngrid * ngrid) if im is None else np.array(im, dtype=np.float64)
        xl = xlambda0 / harm
        self.s = OrcField(nslice, ngrid, first, harm, 1, 0, gridmax, 2 * gridmax / (ngrid - 1),
                          2 * np.pi / xl, xl, slicelength, 0.0, dp(self.re), dp(self.im))

    @property
    def ptr(self):
        self.s.re = dp(self.re)
        self.s.im = dp(self.im)
        return C.byref(self.s)

    @property
    def first(self):
        return self.s.first


def sc_field(beam, ks, gz2, rmax, ngrid, nz, nphi):
    """orc_sc_field: (per-particle push term E_z/511000, per-slice RMS E_z)."""
    S = beam.offsets.size - 1
    ez = np.empty(beam.x.size)
    rms = np.empty(S)
    ef = EField(rmax, ngrid, nz, nphi, 0)
    lib().orc_sc_field(S, ip(beam.offsets), dp(beam.x), dp(beam.y), dp(beam.theta),
                       dp(np.ascontiguousarray(beam.current, dtype=np.float64)), ks, gz2,
                       C.byref(ef), dp(ez), dp(rms))
    return ez, rms


def run(beam, field: OField, steps, *, gamma0, lambda0, solver, sample=1.0, timerun=True,
        periodic=False, nharm=1, do_filter=False, xc=1.0, yc=1.0, sig=1.0, slice_eloss=None,
        efield=None):
    """orc_run on copies of the beam arrays; returns (beam arrays, braw, fraw).
    efield = (rmax, ngrid, nz, nphi) adds the short-range space charge."""
    S = beam.offsets.size - 1
    b = beam.copy()
    nout = 1 + sum(1 for s in steps if s.outstep)
    braw = np.zeros(nout * S * beam_raw_len(nharm))
    fraw = np.zeros(nout * S * FIELD_RAW)
    el = None if slice_eloss is None else dp(np.ascontiguousarray(slice_eloss, dtype=np.float64))
    args = (S, ip(b.offsets), dp(b.x), dp(b.y), dp(b.px), dp(b.py), dp(b.gamma), dp(b.theta),
            gamma0, lambda0, dp(b.current), el, field.ptr, solver, int(do_filter), xc, yc, sig,
            steps, len(steps), sample, int(timerun), int(periodic), nharm, dp(braw), dp(fraw))
    if efield is None:
        n = lib().orc_run(*args)
    else:
        ef = EField(float(efield[0]), int(efield[1]), int(efield[2]), int(efield[3]), 0)
        n = lib().orc_run_sc(*args, C.byref(ef))
    return b, braw.reshape(n, S, -1), fraw.reshape(n, S, FIELD_RAW)
