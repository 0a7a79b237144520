"""CPU-side checks of the C ABI library: it loads, exports every symbol include/*.h
declares, and its host-only entry points (generator, packer, argument checks) behave.
No device compute is called here."""
import ctypes
import os
import re

import numpy as np
import pytest

from conftest import ROOT

bw = pytest.importorskip("biwfa_gpu_pii2025_amd")


def _declared_functions():
    names = set()
    inc = os.path.join(ROOT, "include")
    for fn in os.listdir(inc):
        if not fn.endswith(".h"):
            continue
        text = open(os.path.join(inc, fn)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(bw_[a-z0-9_]+)\s*\(", text):
            names.add(m.group(1))
    return names


def test_library_loads_and_exports_all_declared_symbols():
    L = bw.lib()
    declared = _declared_functions()
    assert len(declared) >= 20
    missing = [n for n in sorted(declared) if not hasattr(L, n)]
    assert not missing, "not exported: %s" % missing


def test_no_oracle_in_product_library():
    """The product library must not contain or link the CPU checkers."""
    data = open(bw.LIB_PATH, "rb").read()
    for sym in (b"bwo_align", b"wfa2ref_align_batch", b"wavefront_align"):
        assert sym not in data


def test_strerror_and_defaults():
    assert "completed" in bw.strerror(0)
    assert "steps" in bw.strerror(-100).lower()
    assert "guard" in bw.strerror(-405)


def test_invalid_penalties_rejected_without_device():
    L = bw.lib()
    a = bw._Attr()
    a.penalties = bw._Pen(0, 0, 6, 2)          # mismatch must be > 0
    assert not L.bw_aligner_new(ctypes.byref(a))
    assert "penalties" in bw.last_error()
    a.penalties = bw._Pen(1, 4, 6, 2)          # match must be <= 0
    assert not L.bw_aligner_new(ctypes.byref(a))
    a.penalties = bw._Pen(0, 4, 40, 2)         # ring deeper than supported
    assert not L.bw_aligner_new(ctypes.byref(a))


def test_generator_deterministic_and_prefix_stable():
    cfg = bw.CONFIGS["C4"]
    a = bw.generate_dataset(cfg, num_pairs=300)
    b = bw.generate_dataset(cfg, num_pairs=300)
    c = bw.generate_dataset(cfg, num_pairs=120)
    assert a.pairs() == b.pairs()
    assert a.pairs()[:120] == c.pairs()     # pair i depends only on (seed, i)
    d = bw.generate_dataset(cfg, num_pairs=50, seed=99)
    assert d.pairs() != a.pairs()[:50]


@pytest.mark.parametrize("name", ["C1", "C2", "C3", "C4", "C5"])
def test_generator_matches_config(name):
    cfg = bw.CONFIGS[name]
    n = {"C1": 50, "C2": 2000, "C3": 5, "C4": 400, "C5": 2}[name]
    ds = bw.generate_dataset(cfg, num_pairs=n)
    assert len(ds) == n
    if cfg.len_min == cfg.len_max:
        assert set(ds.p_len.tolist()) == {cfg.len_min}
    else:
        assert ds.p_len.min() >= cfg.len_min and ds.p_len.max() <= cfg.len_max
        # log-uniform: the median sits near the geometric mean of the range
        gm = np.sqrt(cfg.len_min * cfg.len_max)
        assert 0.5 * gm < np.median(ds.p_len) < 2.0 * gm
    for i in range(min(n, 20)):
        p, t = ds.pair(i)
        assert set(p) <= set("ACGT") and set(t) <= set("ACGT")
    if name == "C2":
        # 3 single-base edits: |t - p| <= 3 and the pairs differ
        assert np.all(np.abs(ds.t_len - ds.p_len) <= 3)


def test_pack_roundtrip():
    rng = np.random.default_rng(1)
    pairs = [("".join(rng.choice(list("ACGT"), int(rng.integers(0, 200)))),
              "".join(rng.choice(list("ACGT"), int(rng.integers(0, 200))))) for _ in range(50)]
    pk = bw.pack_pairs(pairs)
    dec = "ACGT"
    for i, (p, t) in enumerate(pairs):
        for s, woff, n in ((p, pk.p_woff[i], pk.p_len[i]), (t, pk.t_woff[i], pk.t_len[i])):
            out = "".join(dec[(int(pk.words[woff + j // 32]) >> (2 * (j % 32))) & 3] for j in range(n))
            assert out == s
    assert pk.p_woff.min() >= 1
    assert len(pk.words) >= pk.t_woff[-1] + (pk.t_len[-1] + 31) // 32 + 2


def test_pack_rejects_non_acgt():
    with pytest.raises(bw.BiwfaError):
        bw.pack_pairs([("ACGN", "ACGT")])


def test_runs_helpers():
    runs = [(14 << 4) | 7, (1 << 4) | 1, (86 << 4) | 7]
    assert bw.runs_to_cigar(runs) == "14M1I86M"
    assert bw.runs_to_ops(runs) == "M" * 14 + "I" + "M" * 86


def build_c_stub(out_dir):
    """Compiles tests/c_stub/alg_line.c (the integration stub) against the public
    header and links it to the product library; returns the executable's path."""
    import subprocess
    src = os.path.join(ROOT, "tests", "c_stub", "alg_line.c")
    exe = os.path.join(str(out_dir), "alg_line")
    libdir = os.path.dirname(bw.LIB_PATH)
    cmd = ["gcc", "-std=c99", "-Wall", "-Wextra", "-Werror", "-I" + os.path.join(ROOT, "include"), src,
           "-L" + libdir, "-lbiwfa_mi355x", "-Wl,-rpath," + libdir, "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return exe


def test_c_stub_compiles_and_links_against_the_header(tmp_path):
    exe = build_c_stub(tmp_path)
    assert os.access(exe, os.X_OK)


def test_integration_doc_shows_the_stub():
    """The integration guide shows the same C stub the tests compile (tests/c_stub/alg_line.c)."""
    doc = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    src = open(os.path.join(ROOT, "tests", "c_stub", "alg_line.c")).read()
    body = src[src.index("int main"):]
    assert body.strip() in doc


def test_attr_default_and_affine2p_penalty_checks():
    """bw_attr_default fills gap-affine with WFA2's defaults and the 2p penalties with
    align_benchmark's (0,4,6,2,24,1); a gap-affine-2p aligner with bad penalties or
    too deep a ring is refused before any device is touched."""
    L = bw.lib()
    a = bw._Attr()
    L.bw_attr_default(ctypes.byref(a))
    assert (a.penalties.match, a.penalties.mismatch, a.penalties.gap_opening, a.penalties.gap_extension) == \
        (0, 4, 6, 2)
    assert a.distance_metric == bw.GAP_AFFINE
    p2 = a.affine2p_penalties
    assert (p2.match, p2.mismatch, p2.gap_opening1, p2.gap_extension1, p2.gap_opening2, p2.gap_extension2) == \
        (0, 4, 6, 2, 24, 1)
    a.distance_metric = bw.GAP_AFFINE_2P
    a.affine2p_penalties = bw._Pen2p(0, 4, 6, 0, 24, 1)     # e1 must be > 0
    assert not L.bw_aligner_new(ctypes.byref(a))
    assert "2p" in bw.last_error()
    a.affine2p_penalties = bw._Pen2p(0, 4, 6, 2, 80, 1)     # ring deeper than 64 scores
    assert not L.bw_aligner_new(ctypes.byref(a))
    assert "64" in bw.last_error()
    a.distance_metric = 7
    assert not L.bw_aligner_new(ctypes.byref(a))
    assert "metric" in bw.last_error()


def test_ctypes_mirror_matches_the_header(tmp_path):
    """The ctypes structures of the Python front-end have the sizes and field offsets
    the C compiler gives the header's structs."""
    import subprocess
    structs = {"bw_attr_t": (bw._Attr, ["penalties", "scope", "max_alignment_steps", "device", "schedule",
                                         "waves_per_cu", "queue_chunk", "distance_metric", "affine2p_penalties"]),
               "bw_batch_t": (bw._Batch, ["num_pairs", "num_words", "words", "pattern_word_offset",
                                           "pattern_length", "text_word_offset", "text_length"]),
               "bw_result_t": (bw._Result, ["score", "wf_score", "status", "cigar_offset", "cigar_runs",
                                             "cigar_capacity", "cigar_total"]),
               "bw_tuning_t": (bw._Tuning, ["lds_ring_width", "lds_seq_cap", "team_waves", "queue_chunk"]),
               "bw_tier_shape_t": (bw._TierShape, ["lds", "waves_per_aligner", "ring_width", "blocks",
                                                    "waves_per_block", "aligners", "kernel"]),
               "bw_device_result_t": (bw._DevResult, ["score", "wf_score", "status", "num_runs", "runs",
                                                       "run_offset"])}
    lines = ["#include <stdio.h>", "#include <stddef.h>", "#include \"biwfa_mi355x.h\"", "int main(void) {"]
    for name, (_, fields) in structs.items():
        lines.append('  printf("%s %%zu\\n", sizeof(%s));' % (name, name))
        for f in fields:
            lines.append('  printf("%s.%s %%zu\\n", offsetof(%s, %s));' % (name, f, name, f))
    lines += ["  return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    r = subprocess.run(["gcc", "-std=c99", "-I" + os.path.join(ROOT, "include"), str(src), "-o", str(exe)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    got = dict(line.split() for line in subprocess.run([str(exe)], capture_output=True, text=True,
                                                       check=True).stdout.splitlines())
    for name, (cls, fields) in structs.items():
        assert int(got[name]) == ctypes.sizeof(cls), name
        for f in fields:
            assert int(got[name + "." + f]) == getattr(cls, f).offset, name + "." + f
