"""GPU CIGARs of the hand-derived tie cases, through every recording kernel.

The cases of tie_cases.py are decided by one comparison of the backtrace
priority each (pinned on the oracle in test_oracle_ties.py).  A common
prefix of S bases leaves a case's CIGAR as |S|M + its own, so the prefix
length picks the kernel that records the pair (the CIGAR form chooses by
length bucket and capacity); the test checks the kernel that actually ran
through tsu_cigars_kernels.  From 2,000 bases on, the bucket's window is
pinned through the tune table: left to the probe, these near-identical
pairs would all be scored by it and recorded in the narrowest window.
"""
import random
import sys

import pytest

from tie_cases import HOMOPOLYMER_462, TIES_462, TIES_531

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"\n[gpu] {msg}", flush=True)
    sys.stdout.flush()


def prefixed(cases, n, seed):
    rng = random.Random(seed)
    pre = "".join(rng.choice("ACGT") for _ in range(n))
    pairs = [(pre + c[0], pre + c[1]) for c in cases]
    return pairs, [merge_prefix(n, c[3]) for c in cases]


def merge_prefix(n, cigar):
    """|S|M followed by cigar, with a leading M run merged."""
    if n == 0:
        return cigar
    num = ""
    for i, ch in enumerate(cigar):
        if ch.isdigit():
            num += ch
            continue
        if ch == "M":
            return f"{n + int(num)}M" + cigar[i + 1:]
        return f"{n}M" + cigar
    return f"{n}M"


CASES_462 = TIES_462 + HOMOPOLYMER_462

# (name, prefix length, kernel expected to record, pinned window: (bucket bound, diagonals))
PATHS_462 = [
    ("one wave, 1 chunk", 0, "reg1", None),
    ("one wave, 8 chunks", 300, "reg1", None),
    ("one wave, 16 chunks", 700, "reg1", None),
    ("16 chunks x 4 waves", 2000, "ml4", (2048, 4096)),
    ("16 chunks x 8 waves", 5000, "ml8", (8192, 8192)),
    ("16 chunks x 16 waves, doubled offsets", 10000, "ml16", (16383, 16384)),
    ("16 chunks x 16 waves, plain offsets", 20000, "ml16", (32000, 16384)),
    ("LDS kernel, int32 offsets", 33000, "lds", None),
]


@pytest.mark.parametrize("name,n,kernel,pin", PATHS_462, ids=[p[0] for p in PATHS_462])
def test_tie_cigars_462(aligner, name, n, kernel, pin):
    pairs, want = prefixed(CASES_462, n, 11 + n)
    say(f"tie CIGARs 4/6/2, {name}: {len(pairs)} pairs, prefix {n}")
    if pin:
        aligner.set_tune_table([{"waves_per_pair": 1, "capacity": pin[1], "bucket_bound": pin[0]}])
    try:
        got, cig, t = aligner.align_cigar(aligner.batch(pairs))
    finally:
        aligner.set_tune_table([])
    assert t["kernels"] == {kernel}, t["kernels"]
    assert cig == want


def test_tie_cigars_462_probe_placed(aligner):
    """Left to the probe, the prefixed pairs of 2-20 kbp are scored by it
    (their scores are far below its bound) and recorded at their exact
    capacity: by the single-wave recorder, or, beyond the 16,000 bases it
    holds, by the LDS kernel."""
    pairs, want = [], []
    for n in (2000, 5000, 10000, 20000):
        p, w = prefixed(CASES_462, n, 11 + n)
        pairs += p
        want += w
    say(f"tie CIGARs 4/6/2, probe-placed: {len(pairs)} pairs")
    got, cig, t = aligner.align_cigar(aligner.batch(pairs))
    assert t["kernels"] == {"reg1", "lds"}, t["kernels"]
    assert cig == want


@pytest.mark.parametrize("n", [0, 2000])
def test_tie_cigars_462_checkpointed(tsu, n):
    """The LDS kernel's checkpointed backtrace, one score per window."""
    pairs, want = prefixed(CASES_462, n, 23 + n)
    a = tsu.Aligner(4, 6, 2)
    try:
        a.set_cigar_memory(a.CIGAR_CHECKPOINTS, 1)
        say(f"tie CIGARs 4/6/2, checkpoints every score, prefix {n}")
        got, cig, t = a.align_cigar(a.batch(pairs))
        assert t["kernels"] == {"lds_ck"}, t["kernels"]
        assert cig == want
    finally:
        a.close()


@pytest.mark.parametrize("n,checkpoints", [(0, False), (2000, False), (0, True), (2000, True)])
def test_tie_cigars_531(tsu, n, checkpoints):
    """5/3/1 penalties run on the LDS kernel (records, or checkpoints every
    score)."""
    pairs, want = prefixed(TIES_531, n, 31 + n)
    a = tsu.Aligner(5, 3, 1)
    try:
        if checkpoints:
            a.set_cigar_memory(a.CIGAR_CHECKPOINTS, 1)
        say(f"tie CIGARs 5/3/1, prefix {n}, checkpoints {checkpoints}")
        got, cig, t = a.align_cigar(a.batch(pairs))
        assert t["kernels"] == {"lds_ck" if checkpoints else "lds"}, t["kernels"]
        assert cig == want
    finally:
        a.close()
