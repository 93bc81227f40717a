"""ctypes binding of the g4hip C ABI (include/g4hip.h).

The library is the in-tree build genesis4_cuda_port_amd/_lib/libg4hip.so.
Loading it never falls back to anything else: if it is missing, the import of
this module raises.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "_lib" / "libg4hip.so"

G4H_OK = 0
STATUS_NAMES = {
    0: "G4H_OK",
    1: "G4H_ERR_INVALID_ARGUMENT",
    2: "G4H_ERR_NOT_INITIALIZED",
    3: "G4H_ERR_OUT_OF_MEMORY",
    4: "G4H_ERR_DEVICE",
    5: "G4H_ERR_UNSUPPORTED",
    6: "G4H_ERR_NO_DEVICE",
}
FP64, MIXED = 0, 1
ADI, FFT = 0, 1
TUNE_ADI, TUNE_SPLIT, TUNE_BATCH = 0, 1, 2  # kernel families of g4h_context_set_variant

BEAM_STATS = 23
FIELD_RAW = 9
BEAM_Q = 25
FIELD_Q = 9
K_PARTICLE, K_FFT_ROWS, K_FFT_COLS, K_ADI_X, K_ADI_Y, K_OTHER = range(6)
KERNEL_CLASSES = ["particle", "fft_rows", "fft_cols", "adi_x", "adi_y", "other"]


def beam_raw_len(nharm: int) -> int:
    return BEAM_STATS + 2 * nharm


def beam_q_len(nharm: int) -> int:
    return BEAM_Q + 2 * nharm


class G4Error(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {msg}")
        self.status = status


class EField(C.Structure):
    _fields_ = [("rmax", C.c_double), ("ngrid", C.c_int32), ("nz", C.c_int32),
                ("nphi", C.c_int32), ("longrange", C.c_int32)]


class Step(C.Structure):
    _fields_ = [
        ("delz", C.c_double), ("z", C.c_double), ("aw", C.c_double), ("ku", C.c_double),
        ("kx", C.c_double), ("ky", C.c_double), ("gradx", C.c_double), ("grady", C.c_double),
        ("ax", C.c_double), ("ay", C.c_double), ("qf", C.c_double), ("dqx", C.c_double),
        ("dqy", C.c_double), ("cx", C.c_double), ("cy", C.c_double), ("chic_angle", C.c_double),
        ("chic_lb", C.c_double), ("chic_ld", C.c_double), ("chic_lt", C.c_double),
        ("autophase", C.c_double), ("slippage", C.c_double), ("reserved", C.c_double),
        ("helical", C.c_int32), ("in_undulator", C.c_int32), ("marker", C.c_int32),
        ("outstep", C.c_int32),
    ]


assert C.sizeof(Step) == 192


class RunConfig(C.Structure):
    _fields_ = [
        ("lambda0", C.c_double), ("sample", C.c_double), ("timerun", C.c_int32),
        ("periodic", C.c_int32), ("nharm", C.c_int32), ("record_diag", C.c_int32),
    ]


class BeamSpec(C.Structure):
    _fields_ = [
        ("gamma0", C.c_double), ("espread_rel", C.c_double), ("chirp_rel", C.c_double),
        ("emitn_x", C.c_double), ("emitn_y", C.c_double), ("betax", C.c_double),
        ("betay", C.c_double), ("alphax", C.c_double), ("alphay", C.c_double),
        ("xcen", C.c_double), ("ycen", C.c_double), ("pxcen", C.c_double), ("pycen", C.c_double),
        ("slicelength", C.c_double), ("nslice", C.c_int32), ("npart", C.c_int32),
        ("nbins", C.c_int32), ("shotnoise", C.c_int32), ("seed", C.c_uint64),
    ]


class BatchRun(C.Structure):
    _fields_ = [
        ("npart", C.c_int32), ("nbins", C.c_int32), ("ngrid", C.c_int32),
        ("shotnoise", C.c_int32), ("seed", C.c_uint64), ("gridmax", C.c_double),
        ("current", C.c_double), ("gamma0", C.c_double), ("espread_rel", C.c_double),
        ("emitn_x", C.c_double), ("emitn_y", C.c_double), ("betax", C.c_double),
        ("betay", C.c_double), ("alphax", C.c_double), ("alphay", C.c_double),
        ("seed_power", C.c_double), ("seed_w0", C.c_double),
        ("aw", C.POINTER(C.c_double)),
    ]


class Request(C.Structure):
    _fields_ = [
        ("kind", C.c_int32), ("precision", C.c_int32), ("solver", C.c_int32),
        ("nslice", C.c_int32), ("npart", C.c_int32), ("nbins", C.c_int32), ("ngrid", C.c_int32),
        ("shotnoise", C.c_int32), ("nharm", C.c_int32), ("periodic", C.c_int32),
        ("nsteps", C.c_int32), ("seed", C.c_uint64), ("gridmax", C.c_double),
        ("current", C.c_double), ("gamma0", C.c_double), ("espread_rel", C.c_double),
        ("chirp_rel", C.c_double), ("emitn", C.c_double), ("beta", C.c_double),
        ("seed_power", C.c_double), ("seed_w0", C.c_double), ("lambda0", C.c_double),
        ("sample", C.c_double), ("steps", C.POINTER(Step)), ("aw", C.POINTER(C.c_double)),
    ]


class ResultInfo(C.Structure):
    _fields_ = [("status", C.c_int32), ("nrecords", C.c_int32), ("latency_s", C.c_double),
                ("queued_s", C.c_double)]


class ServerStats(C.Structure):
    _fields_ = [("completed", C.c_int64), ("p50_s", C.c_double), ("p99_s", C.c_double),
                ("mean_s", C.c_double), ("particle_slice_steps_per_s", C.c_double),
                ("hbm_in_use", C.c_uint64)]


_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)
_vp = C.c_void_p

# name -> (restype, argtypes)
PROTOTYPES = {
    "g4h_abi_version": (C.c_int, []),
    "g4h_device_count": (C.c_int, []),
    "g4h_last_error": (C.c_char_p, []),
    "g4h_context_create": (C.c_int, [C.c_int, C.POINTER(_vp)]),
    "g4h_context_destroy": (None, [_vp]),
    "g4h_context_synchronize": (C.c_int, [_vp]),
    "g4h_context_memory": (C.c_int, [_vp, C.POINTER(C.c_uint64)]),
    "g4h_context_set_variant": (C.c_int, [_vp, C.c_int, C.c_int]),
    "g4h_context_get_variant": (C.c_int, [_vp, C.c_int, C.POINTER(C.c_int), C.c_int,
                                          C.POINTER(C.c_int)]),
    "g4h_beam_create": (C.c_int, [_vp, C.c_int, C.c_int, _ip, _dp, C.c_double, C.POINTER(_vp)]),
    "g4h_beam_destroy": (None, [_vp]),
    "g4h_beam_upload": (C.c_int, [_vp, _dp, _dp, _dp, _dp, _dp, _dp]),
    "g4h_beam_generate": (C.c_int, [_vp, C.POINTER(BeamSpec), _dp]),
    "g4h_field_fill_planes": (C.c_int, [_vp, _dp, _dp, C.c_int]),
    "g4h_beam_download": (C.c_int, [_vp, _dp, _dp, _dp, _dp, _dp, _dp]),
    "g4h_beam_set_current": (C.c_int, [_vp, _dp]),
    "g4h_beam_info": (C.c_int, [_vp, C.POINTER(C.c_int), C.POINTER(C.c_int64), C.POINTER(C.c_int)]),
    "g4h_field_create": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double,
                                   C.c_double, C.c_int, C.POINTER(_vp)]),
    "g4h_field_destroy": (None, [_vp]),
    "g4h_field_upload": (C.c_int, [_vp, _dp, _dp, C.c_int]),
    "g4h_field_download": (C.c_int, [_vp, _dp, _dp, C.POINTER(C.c_int)]),
    "g4h_field_copy": (C.c_int, [_vp, _vp, _vp]),
    "g4h_field_get_ring": (C.c_int, [_vp, C.POINTER(C.c_int), _dp]),
    "g4h_field_set_ring": (C.c_int, [_vp, C.c_int, C.c_double]),
    "g4h_field_set_enabled": (C.c_int, [_vp, C.c_int]),
    "g4h_field_init_solver": (C.c_int, [_vp, C.c_int, C.c_int, C.c_double, C.c_double,
                                        C.c_double, C.c_int]),
    "g4h_track_transverse": (C.c_int, [_vp, _vp, C.POINTER(Step), C.c_double, C.c_int]),
    "g4h_apply_r56": (C.c_int, [_vp, _vp, C.POINTER(Step), C.c_double]),
    "g4h_beam_advance": (C.c_int, [_vp, _vp, C.POINTER(_vp), C.c_int, C.POINTER(Step),
                                   C.c_double, _dp]),
    "g4h_beam_set_particle_ez": (C.c_int, [_vp, _dp]),
    "g4h_beam_add_energy": (C.c_int, [_vp, _vp, _dp]),
    "g4h_field_track": (C.c_int, [_vp, _vp, _vp, C.POINTER(Step), C.c_double]),
    "g4h_apply_slippage": (C.c_int, [_vp, _vp, C.c_double, C.c_double, C.c_int, C.c_int,
                                     C.POINTER(C.c_int)]),
    "g4h_beam_cell_index": (C.c_int, [_vp, _vp, _vp, _ip]),
    "g4h_diag_beam_raw": (C.c_int, [_vp, _vp, C.c_int, _dp]),
    "g4h_diag_beam_raw_noalloc": (C.c_int, [_vp, _vp, C.c_int, _dp]),
    "g4h_beam_init_efield": (C.c_int, [_vp, C.POINTER(EField), C.c_double]),
    "g4h_beam_sc_field": (C.c_int, [_vp, _dp]),
    "g4h_run_download_sc": (C.c_int, [_vp, _dp]),
    "g4h_diag_field_raw": (C.c_int, [_vp, _vp, _dp]),
    "g4h_diag_finalize_beam": (C.c_int, [C.c_int, C.c_int, _dp, _dp, C.c_double, _dp, _dp]),
    "g4h_diag_finalize_field": (C.c_int, [C.c_int, C.c_double, C.c_double, C.c_double, _dp,
                                          _dp, _dp]),
    "g4h_run_create": (C.c_int, [_vp, _vp, _vp, C.POINTER(Step), C.c_int,
                                 C.POINTER(RunConfig), C.POINTER(_vp)]),
    "g4h_run_destroy": (None, [_vp]),
    "g4h_run_capture": (C.c_int, [_vp]),
    "g4h_run_replay": (C.c_int, [_vp]),
    "g4h_run_advance": (C.c_int, [_vp, C.c_int, C.POINTER(C.c_int)]),
    "g4h_run_records": (C.c_int, [_vp, C.POINTER(C.c_int)]),
    "g4h_run_download_diag": (C.c_int, [_vp, _dp, _dp]),
    "g4h_batch_create": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int, C.POINTER(BatchRun),
                                   C.POINTER(Step), C.c_int, C.c_double, C.c_int,
                                   C.POINTER(_vp)]),
    "g4h_batch_destroy": (None, [_vp]),
    "g4h_batch_advance": (C.c_int, [_vp, C.c_int, C.POINTER(C.c_int)]),
    "g4h_batch_records": (C.c_int, [_vp, C.POINTER(C.c_int)]),
    "g4h_batch_download": (C.c_int, [_vp, C.c_int, _dp, _dp]),
    "g4h_server_create": (C.c_int, [C.c_int, C.c_uint64, C.c_int, C.POINTER(_vp)]),
    "g4h_server_destroy": (None, [_vp]),
    "g4h_server_submit": (C.c_int, [_vp, C.POINTER(Request), C.POINTER(C.c_int64)]),
    "g4h_server_wait": (C.c_int, [_vp, C.c_int64, C.c_double, C.POINTER(ResultInfo)]),
    "g4h_server_result": (C.c_int, [_vp, C.c_int64, _dp, _dp]),
    "g4h_server_release": (C.c_int, [_vp, C.c_int64]),
    "g4h_server_stats": (C.c_int, [_vp, C.POINTER(ServerStats)]),
    "g4h_profile_enable": (C.c_int, [_vp, C.c_int]),
    "g4h_profile_read": (C.c_int, [_vp, C.c_int, C.POINTER(C.c_int64), _dp]),
    "g4h_undulator_fc": (C.c_double, [C.c_double, C.c_int, C.c_int]),
    "g4h_sobol": (C.c_int, [C.c_int, C.c_int64, C.c_int64, _dp]),
    "g4h_inverse_normal": (C.c_double, [C.c_double]),
    "g4h_load_beam": (C.c_int, [C.POINTER(BeamSpec), _dp, _dp, _dp, _dp, _dp, _dp, _dp]),
    "g4h_seed_gaussian": (C.c_int, [C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                                    C.c_double, _dp, _dp]),
    "g4h_lattice_fodo": (C.c_int, [C.c_double, _dp, C.c_int, C.c_int, C.c_int, C.c_int,
                                   C.c_double, C.c_int, C.POINTER(Step), C.c_int,
                                   C.POINTER(C.c_int)]),
    "g4h_lattice_parse": (C.c_int, [C.c_char_p, C.c_char_p, C.c_double, C.c_double, C.c_double,
                                    C.c_int, C.c_double, C.POINTER(Step), C.c_int,
                                    C.POINTER(C.c_int)]),
    "g4h_lattice_match": (C.c_int, [C.POINTER(Step), C.c_int, C.c_double, C.c_double, _dp]),
    "g4h_deck_parse": (C.c_int, [C.c_char_p, C.c_char_p, C.c_int, C.POINTER(C.c_int)]),
    "g4h_deck_run": (C.c_int, [C.c_char_p, C.c_char_p, C.c_int, C.POINTER(C.c_int)]),
}


def load(path: Path | str | None = None) -> C.CDLL:
    p = Path(path) if path else LIB_PATH
    if not p.exists():
        raise ImportError(
            f"{p} is not built; run `python -m genesis4_cuda_port_amd.build` (or "
            "__graft_entry__.build()) first")
    lib = C.CDLL(str(p))
    for name, (res, args) in PROTOTYPES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


_LIB: C.CDLL | None = None


def lib() -> C.CDLL:
    global _LIB
    if _LIB is None:
        _LIB = load()
    return _LIB


def check(status: int) -> None:
    if status != G4H_OK:
        msg = lib().g4h_last_error().decode(errors="replace")
        raise G4Error(status, msg)
