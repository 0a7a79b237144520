import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")
UTEST = os.path.join(GOLDEN, "wfa2_utest")
REFERENCE_WFA2 = "/root/reference/WFA2-lib"

# Penalty sets of WFA2's utest (tests/wfa.utest.sh:47-55): (match, x, o, e)
PENALTY_SETS = {
    "": (0, 4, 6, 2),
    ".p0": (0, 1, 2, 1),
    ".p1": (0, 3, 1, 4),
    ".p2": (0, 5, 3, 2),
    ".p3": (-5, 1, 2, 1),
    ".p4": (-2, 3, 1, 4),
    ".p5": (-3, 5, 3, 2),
}


# Gap-affine-2p penalties of WFA2's utest (align_benchmark's defaults,
# tools/align_benchmark/align_benchmark_params.c:62-69): (match, x, o1, e1, o2, e2)
AFFINE2P = (0, 4, 6, 2, 24, 1)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) device")
    config.addinivalue_line("markers", "slow: CPU checks of several minutes; run with BIWFA_RUN_SLOW=1")


def pytest_collection_modifyitems(config, items):
    if os.environ.get("BIWFA_RUN_SLOW") == "1":
        return
    skip = pytest.mark.skip(reason="slow CPU check (several minutes): set BIWFA_RUN_SLOW=1")
    for item in items:
        if "slow" in item.keywords:
            item.add_marker(skip)


def _ensure_oracle():
    """Builds the C restatement if it is missing (gcc, a second)."""
    lib = os.path.join(ROOT, "oracle", "liboracle.so")
    if not os.path.exists(lib):
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "oracle"], check=True,
                       stdout=subprocess.DEVNULL)
    ref = os.path.join(ROOT, "oracle", "_ref", "libwfa2ref.so")
    if not os.path.exists(ref) and os.path.isdir(REFERENCE_WFA2):
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "ref", "-j4"], check=True,
                       stdout=subprocess.DEVNULL)


@pytest.fixture(scope="session")
def oracle():
    _ensure_oracle()
    from oracle.oracle import Oracle
    return Oracle()


@pytest.fixture(scope="session")
def wfa2ref():
    _ensure_oracle()
    from oracle.oracle import Wfa2Ref
    if not Wfa2Ref.available():
        pytest.skip("oracle/_ref not built (reference sources absent)")
    return Wfa2Ref()


@pytest.fixture(scope="session")
def utest_pairs():
    from oracle.oracle import read_seq_file
    return read_seq_file(os.path.join(UTEST, "wfa.utest.seq"))


def golden_alg(tag, score_only=False):
    from oracle.oracle import read_alg_file
    name = "test.biwfa.%saffine%s.alg" % ("score." if score_only else "", tag)
    return read_alg_file(os.path.join(UTEST, name))


def golden_alg2p(score_only=False):
    """WFA2's BiWFA goldens under gap-affine-2p (tests/wfa.utest.check/test.biwfa[.score].affine2p.alg)."""
    from oracle.oracle import read_alg_file
    name = "test.biwfa.%saffine2p.alg" % ("score." if score_only else "")
    return read_alg_file(os.path.join(UTEST, name))
