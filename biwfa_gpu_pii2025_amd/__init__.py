"""biwfa_gpu_pii2025_amd -- bidirectional wavefront alignment (BiWFA) on one MI355X.

Python front-end over libbiwfa_mi355x.so (C ABI in include/biwfa_mi355x.h).  The
compute runs in the library's gfx950 kernels; this module only packs inputs,
calls the ABI and unpacks results.  If the library is missing, every entry point
raises -- there is no CPU fallback.

Surfaces (mirroring the reference's WFA2-lib API, wavefront/wfa.h:125-216):
  * ``WavefrontAligner``  -- one pair at a time: ``align(pattern, text)``, then
    ``score`` / ``cigar`` / ``cigar_ops`` / ``sam_cigar`` / ``status``.
  * ``align_batch``       -- many pairs, host lists in and out.
  * ``DeviceBatch``       -- packed pairs resident in HBM, aligned repeatedly.
  * ``generate_dataset``  -- the seeded synthetic pair generator.
"""
import ctypes
import os

import numpy as np

from .configs import CONFIGS, DatasetConfig  # noqa: F401

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libbiwfa_mi355x.so")

STATUS_ALG_COMPLETED = 0
STATUS_MAX_STEPS_REACHED = -100
STATUS_OOM = -200
STATUS_UNATTAINABLE = -300
STATUS_DEVICE_CAPACITY = -400
STATUS_HIP_ERROR = -401
STATUS_INVALID_ARG = -402
STATUS_NO_DEVICE = -403
STATUS_BAD_SEQUENCE = -404
STATUS_INTERNAL = -405

SCOPE_ALIGNMENT = 0
SCOPE_SCORE = 1
SCHEDULE_INPUT_ORDER = 0
SCHEDULE_LENGTH_SORTED = 1
SCHEDULE_LEARNED = 2

OP_CHAR = {7: "M", 8: "X", 1: "I", 2: "D"}


class BiwfaError(RuntimeError):
    pass


class _Pen(ctypes.Structure):
    _fields_ = [("match", ctypes.c_int), ("mismatch", ctypes.c_int),
                ("gap_opening", ctypes.c_int), ("gap_extension", ctypes.c_int)]


class _Pen2p(ctypes.Structure):
    _fields_ = [("match", ctypes.c_int), ("mismatch", ctypes.c_int),
                ("gap_opening1", ctypes.c_int), ("gap_extension1", ctypes.c_int),
                ("gap_opening2", ctypes.c_int), ("gap_extension2", ctypes.c_int)]


GAP_AFFINE = 0
GAP_AFFINE_2P = 1


class _Attr(ctypes.Structure):
    _fields_ = [("penalties", _Pen), ("scope", ctypes.c_int),
                ("max_alignment_steps", ctypes.c_int), ("device", ctypes.c_int),
                ("schedule", ctypes.c_int), ("waves_per_cu", ctypes.c_int),
                ("queue_chunk", ctypes.c_int), ("distance_metric", ctypes.c_int),
                ("affine2p_penalties", _Pen2p)]


class _Batch(ctypes.Structure):
    _fields_ = [("num_pairs", ctypes.c_int64), ("num_words", ctypes.c_int64),
                ("words", ctypes.c_void_p), ("pattern_word_offset", ctypes.c_void_p),
                ("pattern_length", ctypes.c_void_p), ("text_word_offset", ctypes.c_void_p),
                ("text_length", ctypes.c_void_p)]


class _Result(ctypes.Structure):
    _fields_ = [("score", ctypes.c_void_p), ("wf_score", ctypes.c_void_p),
                ("status", ctypes.c_void_p), ("cigar_offset", ctypes.c_void_p),
                ("cigar_runs", ctypes.c_void_p), ("cigar_capacity", ctypes.c_int64),
                ("cigar_total", ctypes.c_int64)]


class _DevResult(ctypes.Structure):
    """bw_device_result_t: per-pair outputs in device memory."""
    _fields_ = [("score", ctypes.c_void_p), ("wf_score", ctypes.c_void_p),
                ("status", ctypes.c_void_p), ("num_runs", ctypes.c_void_p),
                ("runs", ctypes.c_void_p), ("run_offset", ctypes.c_void_p)]


class _TierShape(ctypes.Structure):
    """bw_tier_shape_t: launch shape of one tier of the last device-batch alignment."""
    _fields_ = [("lds", ctypes.c_int), ("waves_per_aligner", ctypes.c_int), ("ring_width", ctypes.c_int),
                ("blocks", ctypes.c_int), ("waves_per_block", ctypes.c_int), ("aligners", ctypes.c_int64),
                ("kernel", ctypes.c_char * 96)]


MAX_TIERS = 8
_PROGRESS_FN = ctypes.CFUNCTYPE(None, ctypes.c_char_p, ctypes.c_void_p)


class _Tuning(ctypes.Structure):
    _fields_ = [("lds_ring_width", ctypes.c_int), ("lds_seq_cap", ctypes.c_int),
                ("team_waves", ctypes.c_int * MAX_TIERS), ("queue_chunk", ctypes.c_int * MAX_TIERS)]


class _DatasetParams(ctypes.Structure):
    _fields_ = [("num_pairs", ctypes.c_int64), ("seed", ctypes.c_uint64),
                ("len_min", ctypes.c_int32), ("len_max", ctypes.c_int32),
                ("num_edits", ctypes.c_int32), ("rate_min", ctypes.c_double),
                ("rate_max", ctypes.c_double), ("geometric_runs", ctypes.c_int32),
                ("run_frac", ctypes.c_double), ("run_min", ctypes.c_int32),
                ("run_max", ctypes.c_int32)]


_LIB = None


def lib():
    """Loads libbiwfa_mi355x.so (raises if it has not been built)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise BiwfaError("%s is missing; build it with `python -m biwfa_gpu_pii2025_amd.build`"
                             % LIB_PATH)
        L = ctypes.CDLL(LIB_PATH, mode=ctypes.RTLD_GLOBAL)
        L.bw_aligner_new.restype = ctypes.c_void_p
        L.bw_aligner_new.argtypes = [ctypes.POINTER(_Attr)]
        L.bw_attr_default.argtypes = [ctypes.POINTER(_Attr)]
        L.bw_aligner_delete.argtypes = [ctypes.c_void_p]
        L.bw_align.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_int, ctypes.c_char_p, ctypes.c_int]
        L.bw_align_packed2bits.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_int,
                                           ctypes.c_char_p, ctypes.c_int]
        for name in ("bw_get_score", "bw_get_wf_score", "bw_get_status"):
            getattr(L, name).argtypes = [ctypes.c_void_p]
        L.bw_get_cigar_ops.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p),
                                       ctypes.POINTER(ctypes.c_int)]
        L.bw_get_sam_cigar.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_void_p),
                                       ctypes.POINTER(ctypes.c_int)]
        L.bw_cigar_sprint.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_int]
        L.bw_last_error.restype = ctypes.c_char_p
        L.bw_strerror.restype = ctypes.c_char_p
        L.bw_align_batch.argtypes = [ctypes.c_void_p, ctypes.POINTER(_Batch), ctypes.POINTER(_Result)]
        L.bw_dbatch_upload.restype = ctypes.c_void_p
        L.bw_dbatch_upload.argtypes = [ctypes.c_void_p, ctypes.POINTER(_Batch)]
        L.bw_dbatch_align.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        L.bw_synchronize.argtypes = [ctypes.c_void_p]
        L.bw_dbatch_download.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(_Result)]
        L.bw_dbatch_free.argtypes = [ctypes.c_void_p]
        L.bw_last_timing.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_float),
                                     ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float)]
        L.bw_set_schedule.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]
        L.bw_set_tuning.argtypes = [ctypes.c_void_p, ctypes.POINTER(_Tuning)]
        L.bw_get_tuning.argtypes = [ctypes.c_void_p, ctypes.POINTER(_Tuning)]
        L.bw_last_tier_timing.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int]
        L.bw_last_tier_shape.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(_TierShape)]
        L.bw_last_tier_pairs.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int]
        L.bw_autotune.argtypes = [ctypes.c_void_p, ctypes.POINTER(_Batch), ctypes.c_void_p, ctypes.c_int,
                                  ctypes.POINTER(_Tuning)]
        L.bw_set_progress.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        L.bw_set_graph.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.bw_test_set_ready_guard.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int]
        L.bw_align_device_noalloc.argtypes = [ctypes.c_void_p, ctypes.POINTER(_Batch),
                                              ctypes.POINTER(_DevResult)]
        L.bw_dataset_lengths.argtypes = [ctypes.POINTER(_DatasetParams), ctypes.c_void_p, ctypes.c_void_p]
        L.bw_dataset_generate.argtypes = [ctypes.POINTER(_DatasetParams), ctypes.c_void_p,
                                          ctypes.c_void_p, ctypes.c_void_p]
        L.bw_pack_2bit.argtypes = [ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                   ctypes.c_void_p, ctypes.c_void_p]
        L.bw_seqs_read.restype = ctypes.c_void_p
        L.bw_seqs_read.argtypes = [ctypes.c_char_p, ctypes.c_int]
        L.bw_seqs_count.restype = ctypes.c_int64
        L.bw_seqs_count.argtypes = [ctypes.c_void_p]
        L.bw_seqs_get.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.POINTER(ctypes.c_void_p),
                                  ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_void_p),
                                  ctypes.POINTER(ctypes.c_int)]
        L.bw_seqs_batch.argtypes = [ctypes.c_void_p, ctypes.POINTER(_Batch)]
        L.bw_seqs_free.argtypes = [ctypes.c_void_p]
        L.bw_alg_write.argtypes = [ctypes.c_char_p, ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p,
                                   ctypes.c_void_p, ctypes.c_int]
        L.bw_align_pair.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_char_p, ctypes.c_int,
                                    ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                    ctypes.POINTER(ctypes.c_int32), ctypes.c_char_p, ctypes.c_int,
                                    ctypes.POINTER(ctypes.c_int)]
        _LIB = L
    return _LIB


def last_error():
    return lib().bw_last_error().decode()


def strerror(status):
    return lib().bw_strerror(int(status)).decode()


def runs_to_cigar(runs):
    """uint32 runs -> 'NM..' run-length text (cigar_sprint with print_matches)."""
    return "".join("%d%s" % (int(r) >> 4, OP_CHAR[int(r) & 15]) for r in runs)


def runs_to_ops(runs):
    return "".join(OP_CHAR[int(r) & 15] * (int(r) >> 4) for r in runs)


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


# ---------------------------------------------------------------------------
# packing

class PackedPairs:
    """Pairs packed 2-bit, each sequence word aligned (the bw_batch_t layout)."""

    def __init__(self, words, p_woff, p_len, t_woff, t_len):
        self.words = np.ascontiguousarray(words, np.uint64)
        self.p_woff = np.ascontiguousarray(p_woff, np.int64)
        self.p_len = np.ascontiguousarray(p_len, np.int32)
        self.t_woff = np.ascontiguousarray(t_woff, np.int64)
        self.t_len = np.ascontiguousarray(t_len, np.int32)

    def __len__(self):
        return len(self.p_len)

    def _struct(self):
        b = _Batch()
        b.num_pairs = len(self.p_len)
        b.num_words = len(self.words)
        b.words = _ptr(self.words)
        b.pattern_word_offset = _ptr(self.p_woff)
        b.pattern_length = _ptr(self.p_len)
        b.text_word_offset = _ptr(self.t_woff)
        b.text_length = _ptr(self.t_len)
        return b


def word_layout(p_len, t_len):
    """Word offsets for pattern/text of every pair: one pad word in front, two at the end."""
    p_len = np.asarray(p_len, np.int64)
    t_len = np.asarray(t_len, np.int64)
    pw = (p_len + 31) // 32
    tw = (t_len + 31) // 32
    per = pw + tw
    start = np.zeros(len(per) + 1, np.int64)
    np.cumsum(per, out=start[1:])
    p_woff = 1 + start[:-1]
    t_woff = p_woff + pw
    nwords = int(1 + start[-1] + 2)
    return p_woff, t_woff, nwords


def pack_ascii(ascii_buf, p_off, p_len, t_off, t_len):
    """Packs ASCII sequences (concatenated bytes + offsets) into PackedPairs."""
    n = len(p_len)
    p_woff, t_woff, nwords = word_layout(p_len, t_len)
    words = np.zeros(nwords, np.uint64)
    offs = np.empty(2 * n, np.int64)
    lens = np.empty(2 * n, np.int32)
    woffs = np.empty(2 * n, np.int64)
    offs[0::2], offs[1::2] = p_off, t_off
    lens[0::2], lens[1::2] = p_len, t_len
    woffs[0::2], woffs[1::2] = p_woff, t_woff
    buf = np.frombuffer(ascii_buf, np.uint8) if not isinstance(ascii_buf, np.ndarray) else ascii_buf
    st = lib().bw_pack_2bit(ctypes.c_int64(2 * n), _ptr(buf), _ptr(offs), _ptr(lens), _ptr(woffs),
                            _ptr(words))
    if st != 0:
        raise BiwfaError("sequence %d holds a byte outside ACGT" % (-(st + 1000) // 2))
    return PackedPairs(words, p_woff, p_len, t_woff, t_len)


def pack_pairs(pairs):
    """[(pattern, text), ...] (str) -> PackedPairs."""
    chunks, p_off, p_len, t_off, t_len, pos = [], [], [], [], [], 0
    for p, t in pairs:
        pb, tb = p.encode(), t.encode()
        p_off.append(pos); p_len.append(len(pb)); chunks.append(pb); pos += len(pb)
        t_off.append(pos); t_len.append(len(tb)); chunks.append(tb); pos += len(tb)
    buf = np.frombuffer(b"".join(chunks) + b"\0", np.uint8)
    return pack_ascii(buf, np.array(p_off, np.int64), np.array(p_len, np.int32),
                      np.array(t_off, np.int64), np.array(t_len, np.int32))


# ---------------------------------------------------------------------------
# sequence files and .alg output

SEQ_FORMAT_AUTO, SEQ_FORMAT_WFA2, SEQ_FORMAT_REPO = 0, 1, 2


class SeqFile:
    """The pairs of a sequence file, read by the library (bw_seqs_read): WFA2's .seq
    format (a pattern line and a text line per pair, each behind a one-character
    marker) or the repo driver's input (header "N P T", then N line pairs)."""

    def __init__(self, path, fmt=SEQ_FORMAT_AUTO):
        self._L = lib()
        self._h = self._L.bw_seqs_read(os.fsencode(path), int(fmt))
        if not self._h:
            raise BiwfaError("bw_seqs_read: " + last_error())

    def __len__(self):
        return int(self._L.bw_seqs_count(self._h))

    def pair(self, i):
        p, t = ctypes.c_void_p(), ctypes.c_void_p()
        pl, tl = ctypes.c_int(), ctypes.c_int()
        if self._L.bw_seqs_get(self._h, int(i), ctypes.byref(p), ctypes.byref(pl), ctypes.byref(t),
                               ctypes.byref(tl)) != 0:
            raise IndexError(i)
        return ctypes.string_at(p, pl.value).decode(), ctypes.string_at(t, tl.value).decode()

    def pairs(self):
        return [self.pair(i) for i in range(len(self))]

    def packed(self):
        """The pairs 2-bit packed by the library (copied into a PackedPairs)."""
        b = _Batch()
        st = self._L.bw_seqs_batch(self._h, ctypes.byref(b))
        if st != 0:
            raise BiwfaError("bw_seqs_batch: " + last_error())
        n, nw = b.num_pairs, b.num_words

        def arr(ptr, ct, count, dt):
            if count == 0:
                return np.zeros(0, dt)
            return np.ctypeslib.as_array(ctypes.cast(ptr, ctypes.POINTER(ct)), shape=(count,)).astype(dt)
        return PackedPairs(arr(b.words, ctypes.c_uint64, nw, np.uint64),
                           arr(b.pattern_word_offset, ctypes.c_int64, n, np.int64),
                           arr(b.pattern_length, ctypes.c_int32, n, np.int32),
                           arr(b.text_word_offset, ctypes.c_int64, n, np.int64),
                           arr(b.text_length, ctypes.c_int32, n, np.int32))

    def close(self):
        if getattr(self, "_h", None):
            self._L.bw_seqs_free(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def write_alg(path, result, score_only=False):
    """Writes align_benchmark's lite .alg file (bw_alg_write) from a BatchResult."""
    n = len(result.score)
    st = lib().bw_alg_write(os.fsencode(path), n, _ptr(result.score),
                            None if score_only else _ptr(result.cigar_offset),
                            None if score_only else _ptr(result.cigar_runs), int(bool(score_only)))
    if st != 0:
        raise BiwfaError("bw_alg_write: " + last_error())


# ---------------------------------------------------------------------------
# dataset generator

class Dataset:
    """ASCII pairs from the seeded generator (concatenated buffer + offsets)."""

    def __init__(self, seq, p_off, p_len, t_off, t_len):
        self.seq, self.p_off, self.p_len, self.t_off, self.t_len = seq, p_off, p_len, t_off, t_len

    def __len__(self):
        return len(self.p_len)

    def pair(self, i):
        p = self.seq[self.p_off[i]:self.p_off[i] + self.p_len[i]].tobytes().decode()
        t = self.seq[self.t_off[i]:self.t_off[i] + self.t_len[i]].tobytes().decode()
        return p, t

    def pairs(self, idx=None):
        idx = range(len(self)) if idx is None else idx
        return [self.pair(i) for i in idx]

    def pack(self):
        return pack_ascii(self.seq, self.p_off, self.p_len, self.t_off, self.t_len)

    def subset(self, idx):
        return Dataset.from_pairs(self.pairs(idx))

    @staticmethod
    def from_pairs(pairs):
        chunks, p_off, p_len, t_off, t_len, pos = [], [], [], [], [], 0
        for p, t in pairs:
            pb, tb = p.encode(), t.encode()
            p_off.append(pos); p_len.append(len(pb)); chunks.append(pb); pos += len(pb)
            t_off.append(pos); t_len.append(len(tb)); chunks.append(tb); pos += len(tb)
        seq = np.frombuffer(b"".join(chunks) + b"\0", np.uint8).copy()
        return Dataset(seq, np.array(p_off, np.int64), np.array(p_len, np.int32),
                       np.array(t_off, np.int64), np.array(t_len, np.int32))


def _params(cfg, num_pairs=None, seed=None):
    p = _DatasetParams()
    p.num_pairs = int(cfg.num_pairs if num_pairs is None else num_pairs)
    p.seed = int(cfg.seed if seed is None else seed)
    p.len_min, p.len_max = int(cfg.len_min), int(cfg.len_max)
    p.num_edits = int(cfg.num_edits)
    p.rate_min, p.rate_max = float(cfg.rate_min), float(cfg.rate_max)
    p.geometric_runs = int(cfg.geometric_runs)
    p.run_frac = float(cfg.run_frac)
    p.run_min, p.run_max = int(cfg.run_min), int(cfg.run_max)
    return p


def generate_dataset(cfg, num_pairs=None, seed=None):
    """Generates the pairs of a DatasetConfig (optionally only the first num_pairs)."""
    prm = _params(cfg, num_pairs, seed)
    n = prm.num_pairs
    p_len = np.zeros(n, np.int32)
    t_len = np.zeros(n, np.int32)
    st = lib().bw_dataset_lengths(ctypes.byref(prm), _ptr(p_len), _ptr(t_len))
    if st != 0:
        raise BiwfaError("bw_dataset_lengths failed: %d" % st)
    per = p_len.astype(np.int64) + t_len
    start = np.zeros(n + 1, np.int64)
    np.cumsum(per, out=start[1:])
    p_off = start[:-1].copy()
    t_off = p_off + p_len
    seq = np.zeros(int(start[-1]) + 1, np.uint8)
    st = lib().bw_dataset_generate(ctypes.byref(prm), _ptr(p_off), _ptr(t_off), _ptr(seq))
    if st != 0:
        raise BiwfaError("bw_dataset_generate failed: %d" % st)
    return Dataset(seq, p_off, p_len, t_off, t_len)


# ---------------------------------------------------------------------------
# aligner

class BatchResult:
    def __init__(self, n, cap):
        self.score = np.zeros(n, np.int32)
        self.wf_score = np.zeros(n, np.int32)
        self.status = np.zeros(n, np.int32)
        self.cigar_offset = np.zeros(n + 1, np.int64)
        self.cigar_runs = np.zeros(max(1, cap), np.uint32)

    def _struct(self):
        r = _Result()
        r.score, r.wf_score, r.status = _ptr(self.score), _ptr(self.wf_score), _ptr(self.status)
        r.cigar_offset, r.cigar_runs = _ptr(self.cigar_offset), _ptr(self.cigar_runs)
        r.cigar_capacity = len(self.cigar_runs)
        return r

    def runs(self, i):
        return self.cigar_runs[self.cigar_offset[i]:self.cigar_offset[i + 1]]

    def cigar(self, i):
        return runs_to_cigar(self.runs(i))

    def __len__(self):
        return len(self.score)


class WavefrontAligner:
    """A gfx950 BiWFA aligner (the wavefront_aligner_t of WFA2-lib, wavefront/wfa.h:90-118).

    penalties: (match, x, o, e) for gap-affine, or (match, x, o1, e1, o2, e2) for
    gap-affine-2p (WFA2's distance_metric gap_affine_2p)."""

    def __init__(self, penalties=(0, 4, 6, 2), score_only=False, max_steps=0, device=0,
                 schedule=SCHEDULE_LENGTH_SORTED, waves_per_cu=0, queue_chunk=0):
        a = _Attr()
        lib().bw_attr_default(ctypes.byref(a))
        if len(penalties) == 6:
            a.distance_metric = GAP_AFFINE_2P
            a.affine2p_penalties = _Pen2p(*[int(v) for v in penalties])
        elif len(penalties) == 4:
            a.penalties = _Pen(*[int(v) for v in penalties])
        else:
            raise ValueError("penalties: (match, x, o, e) or (match, x, o1, e1, o2, e2)")
        a.scope = SCOPE_SCORE if score_only else SCOPE_ALIGNMENT
        a.max_alignment_steps = int(max_steps)
        a.device = int(device)
        a.schedule = int(schedule)
        a.waves_per_cu = int(waves_per_cu)
        a.queue_chunk = int(queue_chunk)
        self._L = lib()
        self._h = self._L.bw_aligner_new(ctypes.byref(a))
        if not self._h:
            raise BiwfaError("bw_aligner_new: " + last_error())
        self.penalties = tuple(penalties)
        self.score_only = score_only

    def close(self):
        if getattr(self, "_h", None):
            self._L.bw_aligner_delete(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- single pair (wavefront_align) --
    def align(self, pattern, text):
        pb = pattern.encode() if isinstance(pattern, str) else bytes(pattern)
        tb = text.encode() if isinstance(text, str) else bytes(text)
        return self._L.bw_align(self._h, pb, len(pb), tb, len(tb))

    def align_packed2bits(self, pattern_bytes, plen, text_bytes, tlen):
        return self._L.bw_align_packed2bits(self._h, bytes(pattern_bytes), int(plen),
                                            bytes(text_bytes), int(tlen))

    @property
    def status(self):
        return self._L.bw_get_status(self._h)

    @property
    def score(self):
        return self._L.bw_get_score(self._h)

    @property
    def wf_score(self):
        return self._L.bw_get_wf_score(self._h)

    @property
    def cigar_ops(self):
        p, n = ctypes.c_void_p(), ctypes.c_int()
        self._L.bw_get_cigar_ops(self._h, ctypes.byref(p), ctypes.byref(n))
        return ctypes.string_at(p, n.value).decode() if n.value else ""

    def sam_cigar(self, show_mismatches=True):
        p, n = ctypes.c_void_p(), ctypes.c_int()
        self._L.bw_get_sam_cigar(self._h, int(show_mismatches), ctypes.byref(p), ctypes.byref(n))
        if n.value == 0:
            return np.zeros(0, np.uint32)
        return np.ctypeslib.as_array(ctypes.cast(p, ctypes.POINTER(ctypes.c_uint32)), (n.value,)).copy()

    @property
    def cigar(self):
        return runs_to_cigar(self.sam_cigar(True))

    # -- batches --
    def align_batch(self, packed):
        """PackedPairs (host) -> BatchResult (host)."""
        n = len(packed)
        cap = int(np.sum(packed.p_len.astype(np.int64) + packed.t_len)) + n
        res = BatchResult(n, cap)
        b = packed._struct()
        r = res._struct()
        st = self._L.bw_align_batch(self._h, ctypes.byref(b), ctypes.byref(r))
        if st != 0:
            raise BiwfaError("bw_align_batch: %s (%s)" % (strerror(st), last_error()))
        res.cigar_runs = res.cigar_runs[:r.cigar_total]
        return res

    def upload(self, packed):
        return DeviceBatch(self, packed)

    # -- graph mode: replay a captured launch sequence (speed only; results never change) --
    @property
    def graph(self):
        return bool(getattr(self, "_graph", False))

    @graph.setter
    def graph(self, enable):
        st = self._L.bw_set_graph(self._h, 1 if enable else 0)
        if st != 0:
            raise BiwfaError("bw_set_graph failed (%d)" % st)
        self._graph = bool(enable)

    # -- failure-path testing (bw_test_set_ready_guard) --
    def set_ready_guard(self, ticks=0, withhold=False):
        """Bounds the wait for a claimed shared task's ready flag (ticks of the 100 MHz
        device clock; 0 = the 60 s default) and optionally withholds the flags, so
        that every such wait gives up (the tier then ends with BW_STATUS_INTERNAL)."""
        st = self._L.bw_test_set_ready_guard(self._h, int(ticks), 1 if withhold else 0)
        if st != 0:
            raise BiwfaError("bw_test_set_ready_guard failed (%d)" % st)

    # -- launch tuning (speed only; results never change) --
    @property
    def tuning(self):
        """dict(lds_ring_width, lds_seq_cap, team_waves[8], queue_chunk[8]); zeros mean
        the built-in rule.  Applies to batches uploaded after it is set."""
        t = _Tuning()
        self._L.bw_get_tuning(self._h, ctypes.byref(t))
        return {"lds_ring_width": t.lds_ring_width, "lds_seq_cap": t.lds_seq_cap,
                "team_waves": list(t.team_waves), "queue_chunk": list(t.queue_chunk)}

    @tuning.setter
    def tuning(self, d):
        t = _Tuning()
        t.lds_ring_width = int(d.get("lds_ring_width", 0))
        t.lds_seq_cap = int(d.get("lds_seq_cap", 0))
        for i, v in enumerate(list(d.get("team_waves", []))[:MAX_TIERS]):
            t.team_waves[i] = int(v)
        for i, v in enumerate(list(d.get("queue_chunk", []))[:MAX_TIERS]):
            t.queue_chunk[i] = int(v)
        st = self._L.bw_set_tuning(self._h, ctypes.byref(t))
        if st != 0:
            raise BiwfaError("bw_set_tuning: " + last_error())

    def autotune(self, packed, repeats=2, predicted_score=None, progress=None):
        """Times the sample under candidate tunings (bw_autotune); keeps and returns the best.
        With predicted scores (one per sample pair) the sample runs under the learned schedule.
        `progress(str)` receives one line per timed run and one per candidate."""
        cb = None
        if progress is not None:
            cb = _PROGRESS_FN(lambda msg, user: progress(msg.decode()))
        self._L.bw_set_progress(self._h, cb, None)
        b = packed._struct()
        t = _Tuning()
        pred = None
        if predicted_score is not None:
            pred = np.ascontiguousarray(predicted_score, np.float32)
            if len(pred) != len(packed):
                raise BiwfaError("one predicted score per sample pair is needed")
        st = self._L.bw_autotune(self._h, ctypes.byref(b), _ptr(pred), int(repeats), ctypes.byref(t))
        self._L.bw_set_progress(self._h, None, None)
        if st != 0:
            raise BiwfaError("bw_autotune: %s (%s)" % (strerror(st), last_error()))
        return self.tuning

    def tier_timing(self):
        """Per-tier milliseconds of the last device-batch alignment (LDS tier first)."""
        ms = np.zeros(MAX_TIERS, np.float32)
        n = self._L.bw_last_tier_timing(self._h, _ptr(ms), MAX_TIERS)
        return ms[:max(0, n)].tolist()

    def tier_pairs(self):
        """Pairs each tier took on in the last device-batch alignment (re-runs included)."""
        cnt = np.zeros(MAX_TIERS, np.int32)
        n = self._L.bw_last_tier_pairs(self._h, _ptr(cnt), MAX_TIERS)
        return cnt[:max(0, n)].tolist()

    def tier_shapes(self):
        """Launch shape of each tier of the last device-batch alignment: dicts of lds,
        waves_per_aligner, ring_width, blocks, waves_per_block, aligners, kernel."""
        out = []
        for t in range(MAX_TIERS):
            sh = _TierShape()
            if self._L.bw_last_tier_shape(self._h, t, ctypes.byref(sh)) != 0:
                break
            out.append({"lds": bool(sh.lds), "waves_per_aligner": sh.waves_per_aligner,
                        "ring_width": sh.ring_width, "blocks": sh.blocks,
                        "waves_per_block": sh.waves_per_block, "aligners": int(sh.aligners),
                        "kernel": sh.kernel.decode()})
        return out


class DeviceBatch:
    """A packed batch resident on the aligner's device (bw_dbatch_t)."""

    def __init__(self, aligner, packed):
        self.aligner = aligner
        self.n = len(packed)
        self.cap = int(np.sum(packed.p_len.astype(np.int64) + packed.t_len)) + self.n
        b = packed._struct()
        self._h = aligner._L.bw_dbatch_upload(aligner._h, ctypes.byref(b))
        if not self._h:
            raise BiwfaError("bw_dbatch_upload: " + last_error())

    def set_schedule(self, schedule, predicted_score=None):
        """Pair order (and starting workspace tier) of later align() calls.  The
        learned schedule takes one predicted alignment score per pair; it never
        changes any result, only where and when each pair runs."""
        pred = None
        if predicted_score is not None:
            pred = np.ascontiguousarray(predicted_score, np.float32)
            if len(pred) != self.n:
                raise BiwfaError("one predicted score per pair is needed")
        st = self.aligner._L.bw_set_schedule(self.aligner._h, self._h, int(schedule), _ptr(pred))
        if st != 0:
            raise BiwfaError("bw_set_schedule: " + last_error())

    def align(self, sync=True):
        st = self.aligner._L.bw_dbatch_align(self.aligner._h, self._h)
        if st != 0:
            raise BiwfaError("bw_dbatch_align: " + last_error())
        if sync:
            self.synchronize()

    def synchronize(self):
        st = self.aligner._L.bw_synchronize(self.aligner._h)
        if st != 0:
            raise BiwfaError("bw_synchronize: " + last_error())

    def timing(self):
        """(total_ms, main_kernel_ms, spill_kernel_ms) of the last align(), from HIP events."""
        t, m, s = ctypes.c_float(), ctypes.c_float(), ctypes.c_float()
        self.aligner._L.bw_last_timing(self.aligner._h, ctypes.byref(t), ctypes.byref(m), ctypes.byref(s))
        return t.value, m.value, s.value

    def download(self):
        res = BatchResult(self.n, self.cap)
        r = res._struct()
        st = self.aligner._L.bw_dbatch_download(self.aligner._h, self._h, ctypes.byref(r))
        if st != 0:
            raise BiwfaError("bw_dbatch_download: " + last_error())
        res.cigar_runs = res.cigar_runs[:r.cigar_total]
        return res

    def free(self):
        if getattr(self, "_h", None):
            self.aligner._L.bw_dbatch_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def align_pair(pattern, text, x=4, o=6, e=2):
    """The repo-style align-a-pair entry (bw_align_pair): returns (score, ops)."""
    pb, tb = pattern.encode(), text.encode()
    cap = len(pb) + len(tb) + 1
    ops = ctypes.create_string_buffer(cap)
    score, n = ctypes.c_int32(), ctypes.c_int()
    st = lib().bw_align_pair(pb, len(pb), tb, len(tb), x, o, e, ctypes.byref(score), ops, cap,
                             ctypes.byref(n))
    if st != 0:
        raise BiwfaError("bw_align_pair: %s" % strerror(st))
    return score.value, ops.raw[:n.value].decode()
