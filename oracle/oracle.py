"""ctypes front-ends for the CPU checkers under oracle/.

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and by
bench.py's cpu_baseline leg, always as the checker or the timed CPU baseline,
never as the product path.

* ``Oracle``  -- the C restatement of WFA2-lib's BiWFA (oracle/biwfa_oracle.c).
* ``Wfa2Ref`` -- the vendored WFA2-lib v2.3 compiled from the reference sources
  (oracle/_ref/libwfa2ref*.so, built by oracle/Makefile).

Both return, per pair, the .alg score column and the CIGAR as uint32 runs
``(len << 4) | op`` with op ``=``:7 (match), ``X``:8, ``I``:1, ``D``:2 -- the
encoding of WFA2's cigar_get_CIGAR(show_mismatches=true) (alignment/cigar.c:181-240).
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF_DIR = os.path.join(HERE, "_ref")

OP_CODE_TO_CHAR = {7: "M", 8: "X", 1: "I", 2: "D"}
OP_CHAR_TO_CODE = {"M": 7, "X": 8, "I": 1, "D": 2}


def runs_to_cigar(runs):
    """uint32 runs -> run-length CIGAR string as WFA2's cigar_sprint(print_matches=true)
    writes it (alignment/cigar.c:706-740), e.g. '14M1I86M'."""
    return "".join("%d%s" % (int(r) >> 4, OP_CODE_TO_CHAR[int(r) & 15]) for r in runs)


def cigar_to_runs(cigar):
    out, num = [], 0
    for ch in cigar:
        if ch.isdigit():
            num = num * 10 + ord(ch) - 48
        else:
            out.append((num << 4) | OP_CHAR_TO_CODE[ch])
            num = 0
    return out


def read_seq_file(path):
    """'>pattern' / '<text' line pairs (WFA2 tests/wfa.utest.seq format)."""
    pairs, pat = [], None
    with open(path) as fh:
        for line in fh:
            line = line.rstrip("\n")
            if not line:
                continue
            if line[0] == ">":
                pat = line[1:]
            elif line[0] == "<":
                pairs.append((pat, line[1:]))
                pat = None
    return pairs


def read_alg_file(path):
    """'score<TAB>cigar' lines (align_benchmark lite output)."""
    out = []
    with open(path) as fh:
        for line in fh:
            line = line.rstrip("\n")
            if not line:
                continue
            score, cigar = line.split("\t")
            out.append((int(score), cigar))
    return out


def _concat(pairs):
    """ASCII concat buffer plus per-pair offsets/lengths."""
    chunks, p_off, p_len, t_off, t_len, pos = [], [], [], [], [], 0
    for p, t in pairs:
        pb, tb = p.encode(), t.encode()
        p_off.append(pos); p_len.append(len(pb)); chunks.append(pb); pos += len(pb)
        t_off.append(pos); t_len.append(len(tb)); chunks.append(tb); pos += len(tb)
    buf = np.frombuffer(b"".join(chunks) + b"\0" * 64, np.uint8)
    return (buf, np.array(p_off, np.int64), np.array(p_len, np.int32),
            np.array(t_off, np.int64), np.array(t_len, np.int32))


class _Result:
    def __init__(self, n, cap):
        self.alg_score = np.zeros(n, np.int32)
        self.raw_score = np.zeros(n, np.int32)
        self.status = np.zeros(n, np.int32)
        self.runs_off = np.zeros(n + 1, np.int64)
        self.runs_off[1:] = np.cumsum(cap)
        self.runs = np.zeros(max(1, int(self.runs_off[-1])), np.uint32)
        self.nruns = np.zeros(n, np.int32)
        self.elapsed = 0.0

    def cigar_runs(self, i):
        o = self.runs_off[i]
        return self.runs[o:o + self.nruns[i]]

    def cigar(self, i):
        return runs_to_cigar(self.cigar_runs(i))

    def compact(self):
        """(offsets[n+1], runs) with the runs of all pairs back to back."""
        n = len(self.nruns)
        off = np.zeros(n + 1, np.int64)
        np.cumsum(np.maximum(self.nruns, 0), out=off[1:])
        idx = np.repeat(self.runs_off[:-1] - off[:-1], np.maximum(self.nruns, 0)) + np.arange(off[-1])
        return off, self.runs[idx]


def _p(a, ct):
    return a.ctypes.data_as(ctypes.POINTER(ct))


def _prep(pairs_or_buf):
    if isinstance(pairs_or_buf, tuple) and len(pairs_or_buf) == 5:
        buf, p_off, p_len, t_off, t_len = pairs_or_buf
        return (np.ascontiguousarray(buf, np.uint8), np.ascontiguousarray(p_off, np.int64),
                np.ascontiguousarray(p_len, np.int32), np.ascontiguousarray(t_off, np.int64),
                np.ascontiguousarray(t_len, np.int32))
    return _concat(pairs_or_buf)


class _Engine:
    SYMBOL = None

    def align(self, pairs, penalties=(0, 4, 6, 2), score_only=False, threads=0, max_steps=0):
        """pairs: [(pattern, text), ...] or (ascii_buf, p_off, p_len, t_off, t_len).
        penalties: (match, x, o, e) gap-affine, or (match, x, o1, e1, o2, e2) gap-affine-2p."""
        buf, p_off, p_len, t_off, t_len = _prep(pairs)
        n = len(p_len)
        res = _Result(n, (p_len.astype(np.int64) + t_len + 1))
        el = ctypes.c_double(0.0)
        if len(penalties) == 6:
            self._call2p(n, buf, p_off, p_len, t_off, t_len, tuple(int(v) for v in penalties), score_only,
                         max_steps, threads, res, el)
            res.elapsed = el.value
            return res
        m, x, o, e = penalties
        self._call(n, buf, p_off, p_len, t_off, t_len, m, x, o, e, score_only, max_steps,
                   threads, res, el)
        res.elapsed = el.value
        return res


class Wfa2Ref(_Engine):
    """The vendored WFA2-lib, exact BiWFA settings (memory ultralow, heuristic none)."""

    VARIANTS = {"v3": "libwfa2ref.so", "scalar": "libwfa2ref_scalar.so",
                "native": "libwfa2ref_native.so"}

    def __init__(self, variant="v3"):
        path = os.path.join(REF_DIR, self.VARIANTS[variant])
        if not os.path.exists(path):
            raise FileNotFoundError(path + " (build with: make -C oracle ref)")
        self.lib = ctypes.CDLL(path, mode=ctypes.RTLD_LOCAL)
        self.path = path

    @staticmethod
    def available(variant="v3"):
        return os.path.exists(os.path.join(REF_DIR, Wfa2Ref.VARIANTS[variant]))

    def align(self, pairs, penalties=(0, 4, 6, 2), score_only=False, threads=0, max_steps=0,
              show_mismatches=True):
        """As _Engine.align; show_mismatches=False returns the runs of
        cigar_get_CIGAR(cigar, false): X folded into M (op 0)."""
        self.lib.wfa2ref_set_show_mismatches(ctypes.c_int(int(bool(show_mismatches))))
        try:
            return super().align(pairs, penalties, score_only, threads, max_steps)
        finally:
            self.lib.wfa2ref_set_show_mismatches(ctypes.c_int(1))

    def _call(self, n, buf, p_off, p_len, t_off, t_len, m, x, o, e, score_only, max_steps,
              threads, res, el):
        self.lib.wfa2ref_align_batch(
            ctypes.c_int(n), buf.ctypes.data_as(ctypes.c_char_p),
            _p(p_off, ctypes.c_int64), _p(p_len, ctypes.c_int32),
            _p(t_off, ctypes.c_int64), _p(t_len, ctypes.c_int32),
            m, x, o, e, int(score_only), int(max_steps), int(threads),
            _p(res.alg_score, ctypes.c_int32), _p(res.raw_score, ctypes.c_int32),
            _p(res.status, ctypes.c_int32), _p(res.runs, ctypes.c_uint32),
            _p(res.runs_off, ctypes.c_int64), _p(res.nruns, ctypes.c_int32),
            ctypes.byref(el))

    def _call2p(self, n, buf, p_off, p_len, t_off, t_len, pen, score_only, max_steps, threads, res, el):
        self.lib.wfa2ref_align_batch_2p(
            ctypes.c_int(n), buf.ctypes.data_as(ctypes.c_char_p),
            _p(p_off, ctypes.c_int64), _p(p_len, ctypes.c_int32),
            _p(t_off, ctypes.c_int64), _p(t_len, ctypes.c_int32),
            *pen, int(score_only), int(max_steps), int(threads),
            _p(res.alg_score, ctypes.c_int32), _p(res.raw_score, ctypes.c_int32),
            _p(res.status, ctypes.c_int32), _p(res.runs, ctypes.c_uint32),
            _p(res.runs_off, ctypes.c_int64), _p(res.nruns, ctypes.c_int32),
            ctypes.byref(el))


class Oracle(_Engine):
    """The C restatement of WFA2-lib's BiWFA (oracle/biwfa_oracle.c)."""

    def __init__(self, path=None):
        path = path or os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            raise FileNotFoundError(path + " (build with: make -C oracle)")
        self.lib = ctypes.CDLL(path, mode=ctypes.RTLD_LOCAL)
        self.path = path

    @staticmethod
    def available():
        return os.path.exists(os.path.join(HERE, "liboracle.so"))

    def align(self, pairs, penalties=(0, 4, 6, 2), score_only=False, threads=0, max_steps=0,
              counters=False):
        return super().align(pairs, penalties, score_only, threads, max_steps)

    def _call(self, n, buf, p_off, p_len, t_off, t_len, m, x, o, e, score_only, max_steps,
              threads, res, el):
        res.wf_score = res.raw_score
        res.cells = np.zeros(n, np.int64)
        res.words = np.zeros(n, np.int64)
        self.lib.bwo_align_batch(
            ctypes.c_int(n), buf.ctypes.data_as(ctypes.c_char_p),
            _p(p_off, ctypes.c_int64), _p(p_len, ctypes.c_int32),
            _p(t_off, ctypes.c_int64), _p(t_len, ctypes.c_int32),
            m, x, o, e, int(score_only), int(max_steps), int(threads),
            _p(res.alg_score, ctypes.c_int32), _p(res.wf_score, ctypes.c_int32),
            _p(res.status, ctypes.c_int32), _p(res.runs, ctypes.c_uint32),
            _p(res.runs_off, ctypes.c_int64), _p(res.nruns, ctypes.c_int32),
            _p(res.cells, ctypes.c_int64), _p(res.words, ctypes.c_int64),
            ctypes.byref(el))

    def _call2p(self, n, buf, p_off, p_len, t_off, t_len, pen, score_only, max_steps, threads, res, el):
        res.wf_score = res.raw_score
        res.cells = np.zeros(n, np.int64)
        res.words = np.zeros(n, np.int64)
        self.lib.bwo_align_batch_2p(
            ctypes.c_int(n), buf.ctypes.data_as(ctypes.c_char_p),
            _p(p_off, ctypes.c_int64), _p(p_len, ctypes.c_int32),
            _p(t_off, ctypes.c_int64), _p(t_len, ctypes.c_int32),
            *pen, int(score_only), int(max_steps), int(threads),
            _p(res.alg_score, ctypes.c_int32), _p(res.wf_score, ctypes.c_int32),
            _p(res.status, ctypes.c_int32), _p(res.runs, ctypes.c_uint32),
            _p(res.runs_off, ctypes.c_int64), _p(res.nruns, ctypes.c_int32),
            _p(res.cells, ctypes.c_int64), _p(res.words, ctypes.c_int64),
            ctypes.byref(el))
