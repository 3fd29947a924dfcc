"""ctypes binding of the CPU oracle (oracle/abea_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg as the checker.  The product path never imports
this module.

The oracle restates f5c's align(): fill FPGA/code/device/align_swi.cl:230-503,
traceback/QC FPGA/code/host/src/align_swi.cpp:1139-1312, emission
GPU/align_cuda.cl:62-110 with the cached level_log_stdv.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# ORC_LIB names another build of the same sources (the sanitizer build that
# the CPU tests run the oracle's checks under)
LIB_PATH = os.environ.get("ORC_LIB") or os.path.join(HERE, "_build", "liborc.so")

MODE_F5C = 0
MODE_FP64 = 1


class Info(C.Structure):
    _fields_ = [
        ("score", C.c_double),
        ("avg_log_emission", C.c_double),
        ("path_len", C.c_int32),
        ("max_gap", C.c_int32),
        ("spanned", C.c_int32),
        ("qc_pass", C.c_int32),
        ("end_event", C.c_int32),
        ("n_bands", C.c_int32),
        ("fills", C.c_int64),
    ]


class Scalings(C.Structure):
    _fields_ = [("scale", C.c_float), ("shift", C.c_float), ("var", C.c_float),
                ("log_var", C.c_float)]


class BatchOut(C.Structure):
    _fields_ = [("pairs", C.c_void_p), ("path", C.c_void_p), ("from_", C.c_void_p),
                ("pair_ptr", C.c_void_p), ("n_pairs", C.c_void_p), ("path_len", C.c_void_p),
                ("score", C.c_void_p), ("avg", C.c_void_p), ("fills", C.c_void_p)]


# event_alignment_t (GPU/f5c.h:191-207), x86-64 layout, 32 bytes
EVENT_ALIGNMENT_DTYPE = np.dtype({
    "names": ["ref_kmer", "ref_position", "read_idx", "event_idx", "rc", "model_kmer",
              "hmm_state"],
    "formats": ["S7", "<i4", "<i4", "<i4", "u1", "S7", "S1"],
    "offsets": [0, 8, 12, 16, 20, 21, 28], "itemsize": 32})

_lib = None


def build(force: bool = False) -> str:
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library missing: {LIB_PATH} (run make -C oracle)")
        L = C.CDLL(LIB_PATH)
        P = C.c_void_p
        L.orc_align.restype = C.c_int32
        L.orc_align.argtypes = [P, C.c_int32, P, C.c_int32, P, C.c_int32, P, C.c_float,
                                C.c_float, C.c_int, P, P, P, P, P, P, C.POINTER(Info)]
        L.orc_kmer_rank.restype = C.c_uint32
        L.orc_kmer_rank.argtypes = [C.c_char_p]
        L.orc_transition_constants.restype = None
        L.orc_transition_constants.argtypes = [C.c_int32, C.c_int32, P]
        L.orc_emission_f32.restype = C.c_float
        L.orc_emission_f32.argtypes = [C.c_float, P, C.c_float, C.c_float]
        L.orc_path_gradient.restype = None
        L.orc_path_gradient.argtypes = [P, P, C.c_int32, P, P, P, C.c_float, C.c_float,
                                        C.c_double, P, P, P, P]
        L.orc_path_score.restype = C.c_double
        L.orc_path_score.argtypes = [P, P, C.c_int32, C.c_char_p, P, C.c_int32, P, C.c_float,
                                     C.c_float]
        L.orc_postalign.restype = C.c_int32
        L.orc_postalign.argtypes = [P, P, P, P, C.c_int32, P, C.c_int32]
        L.orc_recalibrate_model.restype = C.c_int
        L.orc_recalibrate_model.argtypes = [P, P, P, P, C.c_int32, C.c_int]
        L.orc_estimate_scalings_mom.restype = Scalings
        L.orc_estimate_scalings_mom.argtypes = [P, C.c_int32, P, P, C.c_int32]
        L.orc_signal_prefix.restype = None
        L.orc_signal_prefix.argtypes = [P, C.c_int64, P, P]
        L.orc_tstat.restype = None
        L.orc_tstat.argtypes = [P, P, C.c_int64, C.c_int32, P]
        L.orc_detect_peaks.restype = C.c_int64
        L.orc_detect_peaks.argtypes = [P, P, C.c_int64, P]
        L.orc_getevents.restype = C.c_int64
        L.orc_getevents.argtypes = [P, C.c_int64, P, C.c_int64]
        L.orc_align_batch.restype = C.c_int
        L.orc_align_batch.argtypes = [C.c_int32, P, P, P, P, P, P, P, P, P, C.c_int,
                                      C.POINTER(BatchOut), C.c_int]
        L.orc_fills_from_moves.restype = None
        L.orc_fills_from_moves.argtypes = [C.c_int32, P, P, P, P, P]
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def transition_constants(n_events: int, n_kmers: int) -> np.ndarray:
    out = np.zeros(4, np.float64)
    lib().orc_transition_constants(n_events, n_kmers, _p(out))
    return out


def kmer_rank(s: bytes) -> int:
    return int(lib().orc_kmer_rank(s))


def emission(x: float, model_row: np.ndarray, scale: float, shift: float) -> float:
    m = np.ascontiguousarray(model_row, np.float32)
    return float(lib().orc_emission_f32(x, _p(m), scale, shift))


def align(seq: bytes, ev_mean: np.ndarray, model: np.ndarray, scale: float, shift: float,
          mode: int = MODE_F5C, replay_moves: np.ndarray | None = None, want_trace=False,
          want_endcol=False):
    """Align one read.  Returns a dict with pairs (after QC), path/from (before
    QC), moves, info fields and optionally the reference-layout trace."""
    seq = bytes(seq)
    L = len(seq)
    ev = np.ascontiguousarray(ev_mean, np.float32)
    n_e = int(ev.shape[0])
    model = np.ascontiguousarray(model, np.float32).reshape(4096, 3)
    n_k = max(L - 5, 0)
    n_bands = n_e + n_k + 2 if (n_k > 0 and n_e > 0) else 0
    cap = n_e + n_k
    out = np.zeros((max(cap, 1), 2), np.int32)
    path = np.zeros((max(cap, 1), 2), np.int32)
    frm = np.zeros(max(cap, 1), np.uint8)
    moves = np.zeros(max(n_bands, 1), np.uint8)
    trace = np.zeros(max(n_bands, 1) * 100, np.uint8) if want_trace else None
    endcol = np.zeros(max(n_e, 1), np.float32) if want_endcol else None
    rm = None
    if replay_moves is not None:
        rm = np.ascontiguousarray(replay_moves, np.uint8)
        assert rm.shape[0] >= n_bands
    info = Info()
    sbuf = C.create_string_buffer(seq, L + 1)
    n = lib().orc_align(_p(out), cap, C.cast(sbuf, C.c_void_p), L, _p(ev), n_e, _p(model),
                        float(scale), float(shift), mode, _p(rm), _p(moves), _p(path),
                        _p(frm), _p(trace), _p(endcol), C.byref(info))
    pl = info.path_len
    res = {
        "n_pairs": int(n),
        "pairs": out[:n].copy(),
        "path": path[:pl].copy(),
        "from": frm[:pl].copy(),
        "moves": moves[:n_bands].copy(),
        "score": info.score,
        "avg_log_emission": info.avg_log_emission,
        "path_len": pl,
        "max_gap": info.max_gap,
        "spanned": info.spanned,
        "qc_pass": info.qc_pass,
        "end_event": info.end_event,
        "n_bands": info.n_bands,
        "fills": info.fills,
    }
    if want_trace:
        res["trace"] = trace[: n_bands * 100].reshape(n_bands, 100) if n_bands else trace[:0]
    if want_endcol:
        res["endcol"] = endcol[:n_e]
    return res


def path_gradient(path, frm, seq: bytes, ev_mean, model, scale, shift, weight=1.0,
                  partials: bool = False):
    """FP64 gradient of one read's best-path score along a fixed path.

    partials=False: (g_read[3], g_mu[4096], g_sigma[4096]) with g_sigma the
    total derivative in level_stdv when level_log_stdv = log(level_stdv).
    partials=True: (g_read, g_mu, g_sigma, g_lambda), the partial
    derivatives of the DP's function of (level_mean, level_stdv,
    level_log_stdv), the log column read as its own input."""
    path = np.ascontiguousarray(path, np.int32)
    frm = np.ascontiguousarray(frm, np.uint8)
    ev = np.ascontiguousarray(ev_mean, np.float32)
    model = np.ascontiguousarray(model, np.float32).reshape(4096, 3)
    g_read = np.zeros(3, np.float64)
    g_mu = np.zeros(4096, np.float64)
    g_sg = np.zeros(4096, np.float64)
    g_lam = np.zeros(4096, np.float64)
    sbuf = C.create_string_buffer(bytes(seq), len(seq) + 1)
    lib().orc_path_gradient(_p(path), _p(frm), int(path.shape[0]), C.cast(sbuf, C.c_void_p),
                            _p(ev), _p(model), float(scale), float(shift), float(weight),
                            _p(g_read), _p(g_mu), _p(g_sg), _p(g_lam))
    if partials:
        return g_read, g_mu, g_sg, g_lam
    return g_read, g_mu, g_sg + g_lam / model[:, 1].astype(np.float64)


def path_score(path, frm, seq: bytes, ev_mean, model, scale, shift) -> float:
    """FP64 score of a fixed path (boundary + transitions + emissions + end trim)."""
    path = np.ascontiguousarray(path, np.int32)
    frm = np.ascontiguousarray(frm, np.uint8)
    ev = np.ascontiguousarray(ev_mean, np.float32)
    model = np.ascontiguousarray(model, np.float32).reshape(4096, 3)
    return float(lib().orc_path_score(_p(path), _p(frm), int(path.shape[0]), bytes(seq), _p(ev),
                                      int(ev.shape[0]), _p(model), float(scale), float(shift)))


def align_batch(batch, n_threads: int = 0, mode: int = MODE_F5C, want_path: bool = False,
                want_pairs: bool = False):
    """OpenMP CPU baseline / parity reference over a Batch
    (abea_amd.synth.Batch layout), longest reads first.

    Returns a dict: n_pairs, path_len, score (float64; float-valued in
    MODE_F5C), avg, fills, threads; with want_path the pre-QC path and its
    from codes ("path", "from") and with want_pairs the post-QC pairs, read
    i's at pair_ptr[i]."""
    n = batch.n_reads
    cap = batch.n_events.astype(np.int64) + np.maximum(batch.read_len.astype(np.int64) - 5, 0)
    pair_ptr = np.zeros(n, np.int64)
    if n > 1:
        pair_ptr[1:] = np.cumsum(cap)[:-1]
    tot = int(cap.sum()) + 1
    res = {"n_pairs": np.zeros(n, np.int32), "path_len": np.zeros(n, np.int32),
           "score": np.zeros(n, np.float64), "avg": np.zeros(n, np.float64),
           "fills": np.zeros(n, np.int64), "pair_ptr": pair_ptr}
    if want_path:
        res["path"] = np.zeros((tot, 2), np.int32)
        res["from"] = np.zeros(tot, np.uint8)
    if want_pairs:
        res["pairs"] = np.zeros((tot, 2), np.int32)
    o = BatchOut(_p(res.get("pairs")), _p(res.get("path")), _p(res.get("from")), _p(pair_ptr),
                 _p(res["n_pairs"]), _p(res["path_len"]), _p(res["score"]), _p(res["avg"]),
                 _p(res["fills"]))
    seqs = np.ascontiguousarray(batch.seq, np.uint8)
    res["threads"] = int(lib().orc_align_batch(
        n, _p(seqs), _p(np.ascontiguousarray(batch.seq_ptr, np.int64)),
        _p(np.ascontiguousarray(batch.read_len, np.int32)),
        _p(np.ascontiguousarray(batch.ev_mean, np.float32)),
        _p(np.ascontiguousarray(batch.ev_ptr, np.int64)),
        _p(np.ascontiguousarray(batch.n_events, np.int32)),
        _p(np.ascontiguousarray(batch.model, np.float32)),
        _p(np.ascontiguousarray(batch.scale, np.float32)),
        _p(np.ascontiguousarray(batch.shift, np.float32)), int(mode), C.byref(o),
        int(n_threads)))
    return res


def fills_from_moves(n_events, n_kmers, moves, mv_off) -> np.ndarray:
    """Cells computed per read (the reference's fills counter,
    GPU/align_cuda.cl:559) from band moves in the library's packed layout."""
    ne = np.ascontiguousarray(n_events, np.int32)
    nk = np.ascontiguousarray(n_kmers, np.int32)
    mv = np.ascontiguousarray(moves, np.uint32)
    off = np.ascontiguousarray(mv_off, np.int64)
    out = np.zeros(ne.shape[0], np.int64)
    lib().orc_fills_from_moves(int(ne.shape[0]), _p(ne), _p(nk), _p(mv), _p(off), _p(out))
    return out


def postalign(seq: bytes, pairs: np.ndarray):
    """f5c postalign() on one read's AlignedPairs.  Returns (records
    [EVENT_ALIGNMENT_DTYPE], base_to_event_map int32[n_kmers, 2],
    events_per_base)."""
    seq = bytes(seq)
    n_k = len(seq) - 5
    pairs = np.ascontiguousarray(pairs, np.int32).reshape(-1, 2)
    rec = np.zeros(max(pairs.shape[0], 1), EVENT_ALIGNMENT_DTYPE)
    bmap = np.zeros((max(n_k, 1), 2), np.int32)
    epb = C.c_double(0.0)
    sbuf = C.create_string_buffer(seq, len(seq) + 1)
    n = lib().orc_postalign(_p(rec), _p(bmap), C.byref(epb), C.cast(sbuf, C.c_void_p), n_k,
                            _p(pairs), int(pairs.shape[0]))
    # copy as bytes: a structured copy would not keep the (zeroed) padding
    raw = rec.view(np.uint8).reshape(-1, EVENT_ALIGNMENT_DTYPE.itemsize)[:n].copy()
    return raw.view(EVENT_ALIGNMENT_DTYPE).reshape(-1), bmap[:n_k].copy(), float(epb.value)


def recalibrate(model, ev_mean, scalings, records, scale_var: bool = True):
    """f5c recalibrate_model().  scalings = (scale, shift, var, log_var);
    returns (recalibrated, new scalings as float32[4])."""
    model = np.ascontiguousarray(model, np.float32).reshape(4096, 3)
    ev = np.ascontiguousarray(ev_mean, np.float32)
    rec = np.ascontiguousarray(records, EVENT_ALIGNMENT_DTYPE)
    sc = Scalings(*[float(v) for v in scalings])
    ok = lib().orc_recalibrate_model(_p(model), _p(ev), C.byref(sc), _p(rec), int(rec.shape[0]),
                                     int(bool(scale_var)))
    return bool(ok), np.array([sc.scale, sc.shift, sc.var, sc.log_var], np.float32)


def estimate_scalings_mom(seq: bytes, ev_mean, model):
    """f5c estimate_scalings_using_mom(); returns float32[4] (scale, shift,
    var, log_var)."""
    seq = bytes(seq)
    model = np.ascontiguousarray(model, np.float32).reshape(4096, 3)
    ev = np.ascontiguousarray(ev_mean, np.float32)
    sbuf = C.create_string_buffer(seq, len(seq) + 1)
    sc = lib().orc_estimate_scalings_mom(C.cast(sbuf, C.c_void_p), len(seq), _p(model), _p(ev),
                                         int(ev.shape[0]))
    return np.array([sc.scale, sc.shift, sc.var, sc.log_var], np.float32)


FAILED_CALIBRATION = 0x1  # GPU/f5c.h:50
FAILED_ALIGNMENT = 0x2    # GPU/f5c.h:51
FAILED_QUALITY_CHK = 0x4  # GPU/f5c.h:52
MIN_CALIBRATION_VAR = 2.5
MAX_EVENTS_PER_BASE = 5.0  # f5c's filter of reads with too many stays


def process_read(seq: bytes, ev_mean, model, scalings, scale_var: bool = True):
    """f5c's per-read flow after align(): QC-passed pairs -> postalign ->
    recalibrate_model -> read status flag (FAILED_ALIGNMENT when align's QC
    failed; FAILED_CALIBRATION when recalibration did not run or var >
    2.5; then FAILED_QUALITY_CHK when events_per_base > 5, the filter of
    reads with too many stays that f5c applies after a passed calibration).
    Returns a dict."""
    sc = np.asarray(scalings, np.float32)
    r = align(seq, ev_mean, model, float(sc[0]), float(sc[1]))
    out = {"align": r, "records": np.zeros(0, EVENT_ALIGNMENT_DTYPE), "map": None,
           "events_per_base": 0.0, "scalings": sc.copy(), "flag": 0, "calibrated": False}
    if r["n_pairs"] == 0:
        out["flag"] = FAILED_ALIGNMENT
        return out
    rec, bmap, epb = postalign(seq, r["pairs"])
    ok, new = recalibrate(model, ev_mean, sc, rec, scale_var)
    out.update(records=rec, map=bmap, events_per_base=epb, scalings=new, calibrated=ok)
    if not ok or new[2] > MIN_CALIBRATION_VAR:
        out["flag"] = FAILED_CALIBRATION
    elif epb > MAX_EVENTS_PER_BASE:
        out["flag"] = FAILED_QUALITY_CHK
    return out


# ---- the step before align(): event detection (getevents.c) ----------------
# f5c event_t, GPU/f5c.h:113-121: u64 start, f32 length, mean, stdv, 4 B pad
EVENT_DTYPE = np.dtype({"names": ["start", "length", "mean", "stdv"],
                        "formats": ["<u8", "<f4", "<f4", "<f4"],
                        "offsets": [0, 8, 12, 16], "itemsize": 24})


def signal_prefix(x):
    """Prefix sums (S, Q) of a raw signal and of its float squares, in
    double, n + 1 entries each."""
    x = np.ascontiguousarray(x, np.float32)
    S = np.zeros(x.shape[0] + 1, np.float64)
    Q = np.zeros_like(S)
    lib().orc_signal_prefix(_p(x), int(x.shape[0]), _p(S), _p(Q))
    return S, Q


def tstat(S, Q, w: int):
    """The windowed t-statistic of width w over the prefix sums."""
    S = np.ascontiguousarray(S, np.float64)
    Q = np.ascontiguousarray(Q, np.float64)
    n = S.shape[0] - 1
    t = np.zeros(max(n, 0), np.float32)
    lib().orc_tstat(_p(S), _p(Q), n, int(w), _p(t))
    return t


def detect_peaks(t_short, t_long):
    """Event boundaries of the short/long peak detector, in emission order."""
    a = np.ascontiguousarray(t_short, np.float32)
    b = np.ascontiguousarray(t_long, np.float32)
    pk = np.zeros(max(a.shape[0], 1), np.int64)
    n = lib().orc_detect_peaks(_p(a), _p(b), int(a.shape[0]), _p(pk))
    return pk[:n].copy()


def getevents(raw):
    """f5c getevents() (GPU/f5cmisc.h:44): the event table of a raw signal,
    as EVENT_DTYPE records."""
    raw = np.ascontiguousarray(raw, np.float32)
    n = int(raw.shape[0])
    if n == 0:
        return np.zeros(0, EVENT_DTYPE)
    out = np.zeros(n + 1, EVENT_DTYPE)
    m = lib().orc_getevents(_p(raw), n, out.ctypes.data_as(C.c_void_p), n + 1)
    if m <= 0:
        raise RuntimeError(f"orc_getevents failed ({m})")
    return out[:m].copy()
