"""ctypes view of libabea's C ABI (include/abea.h).

The library is loaded from abea-opencl_amd/lib/libabea.so (built in-tree by
abea-opencl_amd/build.py).  A missing library is an error: there is no CPU
fallback anywhere in this package.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(HERE), "lib", "libabea.so")

ABEA_OK = 0
PLAN_PLAIN = 0
PLAN_BATCHED = 1

P = C.c_void_p


class EventT(C.Structure):  # event_t, GPU/f5c.h:113-121
    _fields_ = [("start", C.c_uint64), ("length", C.c_float), ("mean", C.c_float),
                ("stdv", C.c_float)]


class EventTable(C.Structure):  # event_table, GPU/f5c.h:124-130
    _fields_ = [("n", C.c_size_t), ("start", C.c_size_t), ("end", C.c_size_t),
                ("event", C.POINTER(EventT))]


class ModelT(C.Structure):
    _fields_ = [("level_mean", C.c_float), ("level_stdv", C.c_float),
                ("level_log_stdv", C.c_float)]


class ScalingsT(C.Structure):
    _fields_ = [("scale", C.c_float), ("shift", C.c_float), ("var", C.c_float),
                ("log_var", C.c_float)]


class AlignedPair(C.Structure):
    _fields_ = [("ref_pos", C.c_int32), ("read_pos", C.c_int32)]


class BatchT(C.Structure):
    _fields_ = [("n_reads", C.c_int32), ("seq", P), ("seq_ptr", P), ("read_len", P),
                ("events", P), ("ev_mean", P), ("ev_ptr", P), ("n_events", P),
                ("scalings", P), ("model", P), ("read_len_host", P), ("n_events_host", P),
                ("seq_ptr_host", P), ("ev_ptr_host", P)]


class OutputT(C.Structure):
    _fields_ = [("pairs", P), ("path_from", P), ("n_pairs", P), ("path_len", P), ("score", P),
                ("avg_log_emission", P)]


class IndexPairT(C.Structure):  # index_pair_t, GPU/f5c.h:184-188
    _fields_ = [("start", C.c_int32), ("stop", C.c_int32)]


class EventAlignmentT(C.Structure):  # event_alignment_t, GPU/f5c.h:191-207
    _fields_ = [("ref_kmer", C.c_char * 7), ("ref_position", C.c_int32), ("read_idx", C.c_int32),
                ("event_idx", C.c_int32), ("rc", C.c_bool), ("model_kmer", C.c_char * 7),
                ("hmm_state", C.c_char)]


assert C.sizeof(EventAlignmentT) == 32


class PostalignOutputT(C.Structure):
    _fields_ = [("alignment", P), ("base_to_event_map", P), ("n_alignment", P),
                ("events_per_base", P), ("scalings", P), ("read_stat_flag", P)]


FAILED_CALIBRATION = 0x001
FAILED_ALIGNMENT = 0x002
FAILED_QUALITY_CHK = 0x004
REFINE_MOM_START = 1


class TuningT(C.Structure):
    _fields_ = [("core_waves_per_block", C.c_int32), ("core_waves_per_cu", C.c_int32),
                ("traceback_overlap", C.c_int32), ("use_graph", C.c_int32),
                ("fuse_traceback", C.c_int32)]


# every symbol include/abea.h declares
EXPORTS = [
    "abea_last_error", "abea_align", "abea_align_db", "abea_ctx_create", "abea_ctx_destroy",
    "abea_ctx_stream", "abea_pair_capacity", "abea_pair_offsets", "abea_plan_create",
    "abea_plan_destroy", "abea_plan_forward", "abea_plan_move_words", "abea_plan_get_moves",
    "abea_plan_set_replay", "abea_plan_adjoint", "abea_plan_get_tuning", "abea_plan_set_tuning",
    "abea_plan_autotune", "abea_plan_last_times", "abea_plan_workspace_bytes",
    "abea_plan_refine_step", "abea_plan_refine", "abea_estimate_scalings_using_mom",
    "abea_postalign", "abea_recalibrate_model", "abea_estimate_scalings_db", "abea_postalign_db",
    "abea_kmer_offsets", "abea_plan_postalign", "abea_plan_estimate_scalings",
    "abea_getevents", "abea_getevents_db", "abea_getevents_device",
    "abea_ctx_set_getevents_overlap", "abea_ctx_getevents_reruns", "abea_host_last_times",
    "abea_host_contexts", "abea_align_noalloc", "abea_ctx_set_getevents_chunk",
    "abea_ctx_getevents_chunk",
]

_lib = None


class AbeaError(RuntimeError):
    pass


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise AbeaError(f"libabea is not built: {LIB_PATH} missing "
                        "(run python abea-opencl_amd/build.py)")
    L = C.CDLL(LIB_PATH)
    L.abea_last_error.restype = C.c_char_p
    L.abea_last_error.argtypes = []
    L.abea_align.restype = C.c_int32
    L.abea_align.argtypes = [P, C.c_char_p, C.c_int32, EventTable, P, ScalingsT, C.c_float]
    L.abea_align_noalloc.restype = C.c_int32
    L.abea_align_noalloc.argtypes = [P, C.c_char_p, C.c_int32, EventTable, P, ScalingsT,
                                     C.c_float]
    L.abea_align_db.restype = C.c_int
    L.abea_align_db.argtypes = [C.c_int32, P, P, P, P, P, P, P]
    L.abea_ctx_create.restype = C.c_int
    L.abea_ctx_create.argtypes = [C.POINTER(P), C.c_int, P]
    L.abea_ctx_destroy.restype = None
    L.abea_ctx_destroy.argtypes = [P]
    L.abea_ctx_stream.restype = P
    L.abea_ctx_stream.argtypes = [P]
    L.abea_pair_capacity.restype = C.c_int64
    L.abea_pair_capacity.argtypes = [C.c_int32, P, P]
    L.abea_pair_offsets.restype = None
    L.abea_pair_offsets.argtypes = [C.c_int32, P, P, P]
    L.abea_plan_create.restype = C.c_int
    L.abea_plan_create.argtypes = [P, C.POINTER(BatchT), C.c_int, C.POINTER(P)]
    L.abea_plan_destroy.restype = None
    L.abea_plan_destroy.argtypes = [P]
    L.abea_plan_forward.restype = C.c_int
    L.abea_plan_forward.argtypes = [P, C.POINTER(OutputT)]
    L.abea_plan_move_words.restype = C.c_int64
    L.abea_plan_move_words.argtypes = [P]
    L.abea_plan_get_moves.restype = C.c_int
    L.abea_plan_get_moves.argtypes = [P, P]
    L.abea_plan_set_replay.restype = C.c_int
    L.abea_plan_set_replay.argtypes = [P, P]
    L.abea_plan_adjoint.restype = C.c_int
    L.abea_plan_adjoint.argtypes = [P, C.POINTER(OutputT), P, P, P]
    L.abea_plan_get_tuning.restype = C.c_int
    L.abea_plan_get_tuning.argtypes = [P, C.POINTER(TuningT)]
    L.abea_plan_set_tuning.restype = C.c_int
    L.abea_plan_set_tuning.argtypes = [P, C.POINTER(TuningT)]
    L.abea_plan_autotune.restype = C.c_int
    L.abea_plan_autotune.argtypes = [P, C.POINTER(OutputT), C.c_int, C.POINTER(C.c_double)]
    L.abea_plan_last_times.restype = C.c_int
    L.abea_plan_last_times.argtypes = [P, C.POINTER(C.c_double)]
    L.abea_plan_workspace_bytes.restype = C.c_int64
    L.abea_plan_workspace_bytes.argtypes = [P]
    L.abea_plan_refine_step.restype = C.c_int
    L.abea_plan_refine_step.argtypes = [P, C.POINTER(OutputT), P, C.c_float]
    L.abea_plan_refine.restype = C.c_int
    L.abea_plan_refine.argtypes = [P, C.POINTER(OutputT), C.c_int, C.c_float, C.c_int,
                                   C.POINTER(C.c_double)]
    L.abea_estimate_scalings_using_mom.restype = ScalingsT
    L.abea_estimate_scalings_using_mom.argtypes = [C.c_char_p, C.c_int32, P, EventTable]
    L.abea_postalign.restype = C.c_int32
    L.abea_postalign.argtypes = [P, P, C.POINTER(C.c_double), C.c_char_p, C.c_int32, P,
                                 C.c_int32]
    L.abea_recalibrate_model.restype = C.c_bool
    L.abea_recalibrate_model.argtypes = [P, EventTable, C.POINTER(ScalingsT), P, C.c_int32,
                                         C.c_bool]
    L.abea_estimate_scalings_db.restype = C.c_int
    L.abea_estimate_scalings_db.argtypes = [C.c_int32, P, P, P, P, P]
    L.abea_postalign_db.restype = C.c_int
    L.abea_postalign_db.argtypes = [C.c_int32, P, P, P, P, P, P, P, P, P, P, P, P]
    L.abea_kmer_offsets.restype = None
    L.abea_kmer_offsets.argtypes = [C.c_int32, P, P]
    L.abea_plan_postalign.restype = C.c_int
    L.abea_plan_postalign.argtypes = [P, C.POINTER(OutputT), C.POINTER(PostalignOutputT), C.c_int]
    L.abea_plan_estimate_scalings.restype = C.c_int
    L.abea_plan_estimate_scalings.argtypes = [P, P]
    L.abea_getevents.restype = EventTable
    L.abea_getevents.argtypes = [C.c_size_t, P]
    L.abea_getevents_db.restype = C.c_int
    L.abea_getevents_db.argtypes = [C.c_int32, P, P, P]
    L.abea_getevents_device.restype = C.c_int
    L.abea_getevents_device.argtypes = [P, C.c_int32, P, P, P, P]
    L.abea_ctx_set_getevents_overlap.restype = C.c_int
    L.abea_ctx_set_getevents_overlap.argtypes = [P, C.c_int32]
    L.abea_ctx_getevents_reruns.restype = C.c_int64
    L.abea_ctx_getevents_reruns.argtypes = [P]
    L.abea_ctx_set_getevents_chunk.restype = C.c_int
    L.abea_ctx_set_getevents_chunk.argtypes = [P, C.c_int32]
    L.abea_ctx_getevents_chunk.restype = C.c_int32
    L.abea_ctx_getevents_chunk.argtypes = [P]
    L.abea_host_last_times.restype = None
    L.abea_host_last_times.argtypes = [C.POINTER(C.c_double)]
    L.abea_host_contexts.restype = C.c_int32
    L.abea_host_contexts.argtypes = []
    _lib = L
    return L


def check(status: int, what: str) -> None:
    if status != ABEA_OK:
        msg = lib().abea_last_error().decode(errors="replace")
        raise AbeaError(f"{what}: status {status}: {msg}")
