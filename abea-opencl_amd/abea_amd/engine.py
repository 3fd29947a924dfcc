"""Device-resident ABEA batches over libabea's C ABI.

torch provides device memory and the stream; every computation runs in
libabea's HIP kernels.  Mirrors the reference's batch flow (align_ocl,
GPU/align_cuda.c:391-943): flattened reads/events in, per-read
n_event_align_pairs and AlignedPair lists out.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib
from .synth import Batch

P = C.c_void_p


def _ptr(t: torch.Tensor | None):
    return None if t is None else P(t.data_ptr())


def _nptr(a: np.ndarray | None):
    return None if a is None else a.ctypes.data_as(P)


class Context:
    """One libabea context bound to a device and torch's current stream."""

    def __init__(self, device: int = 0, stream: torch.cuda.Stream | None = None):
        if not torch.cuda.is_available():
            raise _lib.AbeaError("libabea needs a HIP device; none is visible")
        self.device = device
        torch.cuda.set_device(device)
        # a stream of our own (default priority): graphs cannot be captured on
        # the legacy default stream; ordering with torch's current stream is
        # made explicit around every launch (enter/leave)
        self.stream = stream or torch.cuda.Stream(device=device)
        h = P()
        _lib.check(_lib.lib().abea_ctx_create(C.byref(h), device, P(self.stream.cuda_stream)),
                   "abea_ctx_create")
        self.handle = h

    def enter(self):
        self.stream.wait_stream(torch.cuda.current_stream(self.device))

    def leave(self):
        torch.cuda.current_stream(self.device).wait_stream(self.stream)

    def close(self):
        if self.handle:
            _lib.lib().abea_ctx_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


@dataclass
class DeviceBatch:
    """A Batch uploaded to HBM (inputs only)."""
    host: Batch
    seq: torch.Tensor
    seq_ptr: torch.Tensor
    read_len: torch.Tensor
    ev_mean: torch.Tensor | None
    events: torch.Tensor | None     # raw event_t records (uint8 view, 24 B each)
    ev_ptr: torch.Tensor
    n_events: torch.Tensor
    scalings: torch.Tensor          # float32 [n, 4]
    model: torch.Tensor             # float32 [4096, 3]
    read_len_h: np.ndarray
    n_events_h: np.ndarray
    seq_ptr_h: np.ndarray
    ev_ptr_h: np.ndarray

    @staticmethod
    def upload(b: Batch, device: int = 0, aos_events: bool = False) -> "DeviceBatch":
        dev = torch.device("cuda", device)

        def t(a, dtype=None):
            return torch.from_numpy(np.ascontiguousarray(a)).to(dev, non_blocking=False)

        events = None
        ev_mean = None
        if aos_events:
            events = t(b.events_aos().view(np.uint8))
        else:
            ev_mean = t(b.ev_mean.astype(np.float32))
        return DeviceBatch(
            host=b, seq=t(b.seq.astype(np.uint8)), seq_ptr=t(b.seq_ptr.astype(np.int64)),
            read_len=t(b.read_len.astype(np.int32)), ev_mean=ev_mean, events=events,
            ev_ptr=t(b.ev_ptr.astype(np.int64)), n_events=t(b.n_events.astype(np.int32)),
            scalings=t(b.scalings()), model=t(b.model.astype(np.float32)),
            read_len_h=np.ascontiguousarray(b.read_len, np.int32),
            n_events_h=np.ascontiguousarray(b.n_events, np.int32),
            seq_ptr_h=np.ascontiguousarray(b.seq_ptr, np.int64),
            ev_ptr_h=np.ascontiguousarray(b.ev_ptr, np.int64))

    @property
    def n_reads(self) -> int:
        return int(self.read_len_h.shape[0])

    def c_struct(self) -> _lib.BatchT:
        s = _lib.BatchT()
        s.n_reads = self.n_reads
        s.seq = _ptr(self.seq)
        s.seq_ptr = _ptr(self.seq_ptr)
        s.read_len = _ptr(self.read_len)
        s.events = _ptr(self.events)
        s.ev_mean = _ptr(self.ev_mean)
        s.ev_ptr = _ptr(self.ev_ptr)
        s.n_events = _ptr(self.n_events)
        s.scalings = _ptr(self.scalings)
        s.model = _ptr(self.model)
        s.read_len_host = _nptr(self.read_len_h)
        s.n_events_host = _nptr(self.n_events_h)
        s.seq_ptr_host = _nptr(self.seq_ptr_h)
        s.ev_ptr_host = _nptr(self.ev_ptr_h)
        return s


def pair_offsets(read_len: np.ndarray, n_events: np.ndarray) -> np.ndarray:
    cap = n_events.astype(np.int64) + np.maximum(read_len.astype(np.int64) - 5, 0)
    off = np.zeros(cap.shape[0], np.int64)
    if cap.shape[0] > 1:
        off[1:] = np.cumsum(cap)[:-1]
    return off


class Plan:
    """A planned batch with its output buffers (device)."""

    def __init__(self, ctx: Context, batch: DeviceBatch, form: int = _lib.PLAN_BATCHED):
        self.ctx = ctx
        self.batch = batch
        self.form = form
        self._cb = batch.c_struct()  # keep host arrays referenced
        h = P()
        _lib.check(_lib.lib().abea_plan_create(ctx.handle, C.byref(self._cb), form, C.byref(h)),
                   "abea_plan_create")
        self.handle = h
        n = batch.n_reads
        self.pair_off = pair_offsets(batch.read_len_h, batch.n_events_h)
        cap = int(_lib.lib().abea_pair_capacity(n, _nptr(batch.read_len_h),
                                                _nptr(batch.n_events_h)))
        dev = torch.device("cuda", ctx.device)
        self.capacity = cap
        self.pairs = torch.zeros((max(cap, 1), 2), dtype=torch.int32, device=dev)
        self.path_from = torch.zeros(max(cap, 1), dtype=torch.uint8, device=dev)
        self.n_pairs = torch.zeros(max(n, 1), dtype=torch.int32, device=dev)
        self.path_len = torch.zeros(max(n, 1), dtype=torch.int32, device=dev)
        self.score = torch.zeros(max(n, 1), dtype=torch.float32, device=dev)
        self.avg = torch.zeros(max(n, 1), dtype=torch.float64, device=dev)
        self.out = _lib.OutputT(_ptr(self.pairs), _ptr(self.path_from), _ptr(self.n_pairs),
                                _ptr(self.path_len), _ptr(self.score), _ptr(self.avg))
        torch.cuda.synchronize(ctx.device)  # inputs uploaded and outputs cleared

    def close(self):
        if getattr(self, "handle", None):
            _lib.lib().abea_plan_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- runs
    generation = 0  # forwards run so far (the adjoint needs the paths of the latest)

    def set_params(self, scalings: torch.Tensor | None = None, model: torch.Tensor | None = None):
        """Overwrite the batch's per-read scalings ([n, 4] f32: scale, shift,
        var, log_var) and/or pore model ([4096, 3] f32: level_mean,
        level_stdv, level_log_stdv) in place; the next forward uses them."""
        n = self.batch.n_reads
        if scalings is not None:
            if tuple(scalings.shape) != (n, 4):
                raise ValueError(f"scalings must be [{n}, 4], got {tuple(scalings.shape)}")
            self.batch.scalings[:n].copy_(scalings.detach().to(torch.float32))
        if model is not None:
            if tuple(model.shape) != (4096, 3):
                raise ValueError(f"model must be [4096, 3], got {tuple(model.shape)}")
            self.batch.model.copy_(model.detach().to(torch.float32))

    def forward(self):
        self.ctx.enter()
        _lib.check(_lib.lib().abea_plan_forward(self.handle, C.byref(self.out)), "abea_plan_forward")
        self.ctx.leave()
        self.generation += 1

    def adjoint(self, grad_score: torch.Tensor):
        """Gradient of sum_r grad_score[r] * score[r] along the paths of the
        last forward: ([n, 3] f32 d/d(shift, scale, var), [4096, 3] f64
        partial derivatives in (level_mean, level_stdv, level_log_stdv))."""
        n = self.batch.n_reads
        dev = grad_score.device
        gs = grad_score.contiguous().to(torch.float32)
        g_scal = torch.empty((max(n, 1), 3), dtype=torch.float32, device=dev)
        g_model = torch.empty((4096, 3), dtype=torch.float64, device=dev)
        self.ctx.enter()
        _lib.check(_lib.lib().abea_plan_adjoint(self.handle, C.byref(self.out), _ptr(gs),
                                                _ptr(g_scal), _ptr(g_model)), "abea_plan_adjoint")
        self.ctx.leave()
        gs.record_stream(self.ctx.stream)
        return g_scal[:n], g_model

    def refine_step(self, damping: float = 1.0, scalings_out: torch.Tensor | None = None):
        """One Gauss-Newton step on every read's (scale, shift) along the
        paths of the last forward; writes the batch's scalings in place
        unless scalings_out ([n, 4] f32, device) is given."""
        dst = self.batch.scalings if scalings_out is None else scalings_out
        self.ctx.enter()
        _lib.check(_lib.lib().abea_plan_refine_step(self.handle, C.byref(self.out), _ptr(dst),
                                                    float(damping)), "abea_plan_refine_step")
        self.ctx.leave()
        return dst

    def refine(self, steps: int, damping: float = 1.0, mom_start: bool = True) -> float:
        """steps x (forward + refine step), graph-replayed, scalings updated
        in place; with mom_start each read first continues from the better
        scoring of its given and its method-of-moments scalings (one more
        forward).  Returns the loop's device ms per step."""
        ms = C.c_double()
        self.ctx.enter()
        _lib.check(_lib.lib().abea_plan_refine(self.handle, C.byref(self.out), int(steps),
                                               float(damping),
                                               _lib.REFINE_MOM_START if mom_start else 0,
                                               C.byref(ms)), "abea_plan_refine")
        self.ctx.leave()
        self.generation += steps + (1 if mom_start else 0)
        return ms.value

    def postalign(self, scale_var: bool = True, records: bool = True, base_map: bool = True,
                  scalings_out: torch.Tensor | None = None) -> dict:
        """f5c's flow after align() for every read, from the pairs of the last
        forward: postalign, recalibrate_model, read status.  Returns the
        device outputs (records at the pair offsets, map entries at
        kmer_offsets(); recalibrated scalings in a new [n, 4] tensor unless
        scalings_out is given, which may be the batch's own scalings)."""
        n = self.batch.n_reads
        dev = self.pairs.device
        if not hasattr(self, "_post"):
            nk = np.maximum(self.batch.read_len_h.astype(np.int64) - 5, 0)
            self.kmer_off = np.concatenate([[0], np.cumsum(nk)[:-1]]).astype(np.int64) \
                if n else np.zeros(0, np.int64)
            self._post = {
                "alignment": torch.zeros((max(self.capacity, 1), 32), dtype=torch.uint8,
                                         device=dev),
                "base_to_event_map": torch.zeros((max(int(nk.sum()), 1), 2), dtype=torch.int32,
                                                 device=dev),
                "n_alignment": torch.zeros(max(n, 1), dtype=torch.int32, device=dev),
                "events_per_base": torch.zeros(max(n, 1), dtype=torch.float64, device=dev),
                "read_stat_flag": torch.zeros(max(n, 1), dtype=torch.int32, device=dev),
            }
        o = dict(self._post)
        o["scalings"] = (torch.zeros((max(n, 1), 4), dtype=torch.float32, device=dev)
                         if scalings_out is None else scalings_out)
        po = _lib.PostalignOutputT(_ptr(o["alignment"]) if records else None,
                                   _ptr(o["base_to_event_map"]) if base_map else None,
                                   _ptr(o["n_alignment"]), _ptr(o["events_per_base"]),
                                   _ptr(o["scalings"]), _ptr(o["read_stat_flag"]))
        self.ctx.enter()
        _lib.check(_lib.lib().abea_plan_postalign(self.handle, C.byref(self.out), C.byref(po),
                                                  int(bool(scale_var))), "abea_plan_postalign")
        self.ctx.leave()
        return o

    def estimate_scalings(self, out: torch.Tensor | None = None) -> torch.Tensor:
        """f5c estimate_scalings_using_mom() for every read ([n, 4] f32)."""
        n = self.batch.n_reads
        dst = out if out is not None else torch.zeros((max(n, 1), 4), dtype=torch.float32,
                                                      device=self.pairs.device)
        self.ctx.enter()
        _lib.check(_lib.lib().abea_plan_estimate_scalings(self.handle, _ptr(dst)),
                   "abea_plan_estimate_scalings")
        self.ctx.leave()
        return dst

    def times(self) -> list[float]:
        t = (C.c_double * 4)()
        _lib.check(_lib.lib().abea_plan_last_times(self.handle, t), "abea_plan_last_times")
        return [float(x) for x in t]

    def autotune(self, reps: int = 1) -> float:
        ms = C.c_double()
        self.ctx.enter()
        _lib.check(_lib.lib().abea_plan_autotune(self.handle, C.byref(self.out), reps, C.byref(ms)),
                   "abea_plan_autotune")
        self.ctx.leave()
        return ms.value

    def tuning(self) -> dict:
        t = _lib.TuningT()
        _lib.check(_lib.lib().abea_plan_get_tuning(self.handle, C.byref(t)), "get_tuning")
        return {f: getattr(t, f) for f, _ in _lib.TuningT._fields_}

    def set_tuning(self, **kw):
        cur = self.tuning()
        cur.update(kw)
        t = _lib.TuningT(**cur)
        _lib.check(_lib.lib().abea_plan_set_tuning(self.handle, C.byref(t)), "set_tuning")

    def moves(self) -> np.ndarray:
        nw = int(_lib.lib().abea_plan_move_words(self.handle))
        m = np.zeros(max(nw, 1), np.uint32)
        _lib.check(_lib.lib().abea_plan_get_moves(self.handle, _nptr(m)), "get_moves")
        return m[:nw]

    def set_replay(self, moves: np.ndarray | None):
        if moves is None:
            _lib.check(_lib.lib().abea_plan_set_replay(self.handle, None), "set_replay")
            return
        m = np.ascontiguousarray(moves, np.uint32)
        self._replay = m
        _lib.check(_lib.lib().abea_plan_set_replay(self.handle, _nptr(m)), "set_replay")

    def workspace_bytes(self) -> int:
        return int(_lib.lib().abea_plan_workspace_bytes(self.handle))

    # ---- results
    def results(self) -> dict:
        """Copy outputs to host.  pairs_i = pairs[off[i]:off[i]+path_len[i]]."""
        torch.cuda.synchronize(self.ctx.device)
        n = self.batch.n_reads
        return {
            "pairs": self.pairs.cpu().numpy(),
            "path_from": self.path_from.cpu().numpy(),
            "n_pairs": self.n_pairs[:n].cpu().numpy(),
            "path_len": self.path_len[:n].cpu().numpy(),
            "score": self.score[:n].cpu().numpy(),
            "avg": self.avg[:n].cpu().numpy(),
            "pair_off": self.pair_off,
        }


# event_alignment_t (GPU/f5c.h:191-207), x86-64 layout, 32 bytes
EVENT_ALIGNMENT_DTYPE = np.dtype({
    "names": ["ref_kmer", "ref_position", "read_idx", "event_idx", "rc", "model_kmer",
              "hmm_state"],
    "formats": ["S7", "<i4", "<i4", "<i4", "u1", "S7", "S1"],
    "offsets": [0, 8, 12, 16, 20, 21, 28], "itemsize": 32})


def postalign_results(plan: "Plan", o: dict) -> dict:
    """Host copies of Plan.postalign() outputs.  Read i's records are
    records[pair_off[i]:pair_off[i] + n_alignment[i]], its map entries
    base_to_event_map[kmer_off[i]:kmer_off[i] + max(read_len[i] - 5, 0)]."""
    torch.cuda.synchronize(plan.ctx.device)
    n = plan.batch.n_reads
    return {
        "records": o["alignment"].cpu().numpy().view(EVENT_ALIGNMENT_DTYPE).reshape(-1),
        "base_to_event_map": o["base_to_event_map"].cpu().numpy(),
        "n_alignment": o["n_alignment"][:n].cpu().numpy(),
        "events_per_base": o["events_per_base"][:n].cpu().numpy(),
        "scalings": o["scalings"][:n].cpu().numpy(),
        "read_stat_flag": o["read_stat_flag"][:n].cpu().numpy(),
        "pair_off": plan.pair_off,
        "kmer_off": plan.kmer_off,
    }


def move_offsets(read_len: np.ndarray, n_events: np.ndarray) -> np.ndarray:
    """Offsets (uint32 words) of each read's band-move bits in Plan.moves()."""
    nk = read_len.astype(np.int64) - 5
    ok = (nk > 0) & (n_events > 0)
    nb = np.where(ok, n_events.astype(np.int64) + nk + 2, 0)
    words = (nb + 31) // 32
    off = np.zeros(words.shape[0], np.int64)
    if words.shape[0] > 1:
        off[1:] = np.cumsum(words)[:-1]
    return off


def unpack_moves(words: np.ndarray, off: int, n_bands: int) -> np.ndarray:
    w = words[off:off + (n_bands + 31) // 32].astype(np.uint32)
    bits = ((w[:, None] >> np.arange(32, dtype=np.uint32)[None, :]) & 1).astype(np.uint8)
    return bits.reshape(-1)[:n_bands]


def pack_moves(bits_per_read: list, read_len, n_events) -> np.ndarray:
    off = move_offsets(np.asarray(read_len), np.asarray(n_events))
    total = 0
    out_parts = []
    for bits in bits_per_read:
        nb = bits.shape[0]
        words = (nb + 31) // 32
        pad = np.zeros(words * 32, np.uint32)
        pad[:nb] = bits
        out_parts.append((pad.reshape(words, 32) << np.arange(32, dtype=np.uint32)).sum(1)
                         .astype(np.uint32))
        total += words
    assert len(off) == len(out_parts)
    return np.concatenate(out_parts) if out_parts else np.zeros(0, np.uint32)


# f5c event_t as a numpy record (u64 start, f32 length, mean, stdv, 4 B pad)
EVENT_DTYPE = np.dtype({"names": ["start", "length", "mean", "stdv"],
                        "formats": ["<u8", "<f4", "<f4", "<f4"],
                        "offsets": [0, 8, 12, 16], "itemsize": 24})


def getevents(ctx: Context, raw: torch.Tensor, raw_off: np.ndarray,
              events: torch.Tensor | None = None, n_events: torch.Tensor | None = None):
    """f5c getevents() (GPU/f5cmisc.h:44) on signals resident in device memory:
    read i is raw[raw_off[i]:raw_off[i+1]] (float32 pA).  Returns (events,
    n_events): a uint8 tensor of 24-byte event_t records, read i's at record
    raw_off[i] + i, and an int64 count per read (both on the device)."""
    raw_off = np.ascontiguousarray(raw_off, np.int64)
    n = raw_off.shape[0] - 1
    if raw.dtype != torch.float32 or not raw.is_cuda:
        raise _lib.AbeaError("getevents: raw must be a float32 device tensor")
    cap = int(raw_off[-1]) + n
    if events is None:
        events = torch.empty(max(cap, 1) * 24, dtype=torch.uint8, device=raw.device)
    if n_events is None:
        n_events = torch.empty(max(n, 1), dtype=torch.int64, device=raw.device)
    ctx.enter()
    _lib.check(_lib.lib().abea_getevents_device(ctx.handle, n, _nptr(raw_off), _ptr(raw),
                                                _ptr(events), _ptr(n_events)),
               "abea_getevents_device")
    ctx.leave()
    return events, n_events


def getevents_results(events: torch.Tensor, n_events: torch.Tensor, raw_off: np.ndarray) -> list:
    """Per-read EVENT_DTYPE arrays from getevents' device outputs."""
    raw_off = np.asarray(raw_off, np.int64)
    n = raw_off.shape[0] - 1
    rec = events.cpu().numpy().view(EVENT_DTYPE)
    cnt = n_events.cpu().numpy()[:n]
    return [rec[raw_off[i] + i: raw_off[i] + i + cnt[i]].copy() for i in range(n)]
