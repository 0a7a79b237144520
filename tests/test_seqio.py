"""The library's sequence-file reader and .alg writer (host code, no GPU):
WFA2's .seq format and the repo driver's input format in, align_benchmark's lite
.alg lines out, checked against the WFA2 utest files."""
import os

import numpy as np
import pytest

from conftest import UTEST, golden_alg

bw = pytest.importorskip("biwfa_gpu_pii2025_amd")
SEQ = os.path.join(UTEST, "wfa.utest.seq")


def test_reads_the_utest_seq_file_like_the_oracle_reader():
    from oracle.oracle import read_seq_file
    want = read_seq_file(SEQ)
    with bw.SeqFile(SEQ) as f:
        assert len(f) == len(want) == 305
        assert f.pairs() == want
        got = f.packed()
    ref = bw.pack_pairs(want)
    for k in ("words", "p_woff", "p_len", "t_woff", "t_len"):
        assert np.array_equal(getattr(got, k), getattr(ref, k)), k


def test_wfa2_reader_quirks(tmp_path):
    # the first line of a pair is the pattern whatever its marker ('<' first is
    # what generate_dataset writes for a longer pattern); a last line without a
    # newline loses its last base, as align_benchmark_read_input does
    p = tmp_path / "x.seq"
    p.write_bytes(b"<ACGTACGT\n>ACG\n>AC\n<GGTT")
    with bw.SeqFile(str(p)) as f:
        assert f.pairs() == [("ACGTACGT", "ACG"), ("AC", "GGT")]


def test_reads_the_repo_driver_format(tmp_path):
    p = tmp_path / "inputs.txt"
    p.write_text("3 5 6 \nACGTA\nACGTTA\nCCCCC\nCCGCCC\nGATTA\nGATTAC\n")
    with bw.SeqFile(str(p)) as f:          # detected: no '>' / '<' marker
        assert f.pairs() == [("ACGTA", "ACGTTA"), ("CCCCC", "CCGCCC"), ("GATTA", "GATTAC")]
    with bw.SeqFile(str(p), bw.SEQ_FORMAT_REPO) as f:
        assert len(f) == 3
    short = tmp_path / "short.txt"
    short.write_text("2 5 5\nACGTA\nACG\nACGTA\nACGTA\n")
    with pytest.raises(bw.BiwfaError, match="shorter"):
        bw.SeqFile(str(short))
    with pytest.raises(bw.BiwfaError, match="cannot open"):
        bw.SeqFile(str(tmp_path / "missing.seq"))


def test_bad_bases_are_reported_at_packing(tmp_path):
    p = tmp_path / "n.seq"
    p.write_text(">ACGT\n<ACNT\n")
    with bw.SeqFile(str(p)) as f:
        assert f.pairs() == [("ACGT", "ACNT")]
        with pytest.raises(bw.BiwfaError, match="outside ACGT"):
            f.packed()


def _result_from(lines):
    from oracle.oracle import cigar_to_runs
    n = len(lines)
    runs = [cigar_to_runs(c) if c != "-" else [] for _, c in lines]
    res = bw.BatchResult(n, sum(len(r) for r in runs))
    for i, (s, _) in enumerate(lines):
        res.score[i] = s
        res.cigar_offset[i + 1] = res.cigar_offset[i] + len(runs[i])
        res.cigar_runs[res.cigar_offset[i]:res.cigar_offset[i + 1]] = runs[i]
    return res


@pytest.mark.parametrize("tag", ["", ".p0", ".p5"])
def test_alg_writer_reproduces_the_golden_files(tmp_path, tag):
    name = "test.biwfa.affine%s.alg" % tag
    out = tmp_path / name
    bw.write_alg(str(out), _result_from(golden_alg(tag)))
    assert out.read_bytes() == open(os.path.join(UTEST, name), "rb").read()
    name = "test.biwfa.score.affine%s.alg" % tag
    out = tmp_path / name
    bw.write_alg(str(out), _result_from(golden_alg(tag, score_only=True)), score_only=True)
    assert out.read_bytes() == open(os.path.join(UTEST, name), "rb").read()


def test_alg_writer_empty_cigar_line(tmp_path):
    # benchmark_print_output: a full alignment with an empty CIGAR prints -1 and '-'
    res = _result_from([(0, "-"), (-4, "1X")])
    out = tmp_path / "e.alg"
    bw.write_alg(str(out), res)
    assert out.read_text() == "-1\t-\n-4\t1X\n"
