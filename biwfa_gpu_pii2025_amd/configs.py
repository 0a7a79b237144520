"""The five synthetic workloads (x=4, o=6, e=2, match=0) and their generator settings.

C1  plain BiWFA       10,000 pairs, 1,000 bp, 5% edits (X:I:D 1:1:1), seed 1
C2  short reads    1,000,000 pairs,   150 bp, 2% edits (3 per pair),  seed 2
C3  long pairs        10,000 pairs, 10,000 bp, 1,000 edit events, I/D runs Geometric(0.5), seed 3
C4  uneven mix       200,000 pairs, log-uniform 100-50,000 bp, rate U(1%,15%), seed 4
C5  divergent long     1,000 pairs, 100,000 bp, 15,000 events, 10% I/D runs U{10..200}, seed 5
"""
from dataclasses import dataclass, replace


@dataclass(frozen=True)
class DatasetConfig:
    name: str
    num_pairs: int
    seed: int
    len_min: int
    len_max: int
    num_edits: int = -1          # >= 0: fixed events per pair
    rate_min: float = 0.0        # else round(U(rate_min, rate_max) * plen)
    rate_max: float = 0.0
    geometric_runs: int = 0      # single-base I/D events get Geometric(0.5) lengths
    run_frac: float = 0.0        # fraction of events that are I/D runs U{run_min..run_max}
    run_min: int = 1
    run_max: int = 1
    description: str = ""

    def with_pairs(self, n, seed=None):
        return replace(self, num_pairs=int(n), seed=self.seed if seed is None else int(seed))


CONFIGS = {
    "C1": DatasetConfig("C1", 10_000, 1, 1000, 1000, num_edits=50,
                        description="plain BiWFA, 1 kbp, 5% edits"),
    "C2": DatasetConfig("C2", 1_000_000, 2, 150, 150, num_edits=3,
                        description="short reads, 150 bp, 2% edits"),
    "C3": DatasetConfig("C3", 10_000, 3, 10_000, 10_000, num_edits=1000, geometric_runs=1,
                        description="long pairs, 10 kbp, 10% edits, geometric indel runs"),
    "C4": DatasetConfig("C4", 200_000, 4, 100, 50_000, rate_min=0.01, rate_max=0.15,
                        description="uneven mix, log-uniform 100-50,000 bp, 1-15% edits"),
    "C4_TRAIN": DatasetConfig("C4_TRAIN", 50_000, 40, 100, 50_000, rate_min=0.01, rate_max=0.15,
                              description="cost-model training set (C4 distribution)"),
    "C5": DatasetConfig("C5", 1_000, 5, 100_000, 100_000, num_edits=15_000, run_frac=0.1,
                        run_min=10, run_max=200,
                        description="divergent long pairs, 100 kbp, 15% edits, 10% long indels"),
}

PENALTIES = (0, 4, 6, 2)
