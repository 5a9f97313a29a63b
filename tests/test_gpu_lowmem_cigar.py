"""GPU parity of the checkpointed (low-memory) backtrace.

The checkpointed backtrace keeps ring snapshots every K scores and recomputes
one window of K scores at a time with records, instead of records for every
score.  Its CIGARs must equal the oracle's (the same CIGARs the full-record
kernels give) for every K, including K = 1 (a snapshot per score) and K
larger than the score (one window), on short, long, int32-offset and
degenerate pairs; the automatic mode must switch to it when records do not
fit the backtrace budget, and then hold less device memory.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
SEED = 0x5453554E


def say(msg):
    print(f"\n[gpu] {msg}", flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ck_aligner(tsu):
    a = tsu.Aligner(4, 6, 2)
    yield a
    a.close()


def _check_vs_oracle(got, cig, pairs, oracle, pen=(4, 6, 2)):
    want, ops, _ = oracle.wfa_batch(pairs, *pen, cigar=True)
    assert (got == want).all(), np.nonzero(got != want)[0][:10]
    for i, (p, t) in enumerate(pairs):
        assert cig[i] == oracle.rle(ops[i]), i
        assert oracle.rescore(oracle.expand(cig[i]), p, t, *pen) == got[i], i


@pytest.mark.parametrize("every", [0, 1, 3, 7, 100000])
def test_checkpoint_cigars_mixed(ck_aligner, tsu, oracle, every):
    """300 pairs of 50-3000 bp at 0-15%: same scores and CIGARs as the oracle
    for snapshots every 1, 3, 7 scores, the automatic interval, and one
    window covering every score."""
    ds = tsu.Dataset.generate(300, seed=SEED + 9, len_min=50, len_max=3000, err_min=0.0, err_max=0.15)
    pairs = ds.pairs()
    say(f"checkpoint cigars mixed 300, every={every}")
    ck_aligner.set_cigar_memory(ck_aligner.CIGAR_CHECKPOINTS, every)
    got, cig, t = ck_aligner.align_cigar(ck_aligner.batch(ds))
    assert t["checkpoint_launches"] >= 1 and t["record_launches"] == 0
    _check_vs_oracle(got, cig, pairs, oracle)


def test_checkpoint_cigars_fixtures(ck_aligner, tsu):
    """The golden fixtures (empty, one-base, identical, all-mismatch and
    seeded pairs) through the checkpointed backtrace."""
    import json
    import os
    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "pairs_scores_cigars.json")) as f:
        cases = [c for c in json.load(f)["cases"] if tuple(c["penalties"]) == (4, 6, 2)]
    say(f"checkpoint cigars fixtures {len(cases)}")
    ck_aligner.set_cigar_memory(ck_aligner.CIGAR_CHECKPOINTS, 2)
    got, cig, _ = ck_aligner.align_cigar(ck_aligner.batch([(c["pattern"], c["text"]) for c in cases]))
    assert [int(v) for v in got] == [c["score"] for c in cases]
    assert cig == [c["cigar"] for c in cases]


def test_checkpoint_cigars_long_reads(ck_aligner, tsu, oracle):
    """10 kbp at 10% and 12-14 kbp at 12% (capacities of 7k-11k diagonals,
    ring in HBM): checkpointed CIGARs equal the oracle's, and the backtrace
    state is a fraction of what records for every score take."""
    pairs = (tsu.Dataset.generate(3, seed=SEED + 13, len_min=10000, err_min=0.10).pairs() +
             tsu.Dataset.generate(2, seed=SEED + 19, len_min=12000, len_max=14000, err_min=0.12).pairs())
    say("checkpoint cigars long 5 pairs")
    ck_aligner.set_cigar_memory(ck_aligner.CIGAR_RECORDS)
    got_r, cig_r, t_r = ck_aligner.align_cigar(ck_aligner.batch(pairs))
    ck_aligner.set_cigar_memory(ck_aligner.CIGAR_CHECKPOINTS)
    got, cig, t = ck_aligner.align_cigar(ck_aligner.batch(pairs))
    say(f"state bytes: records {t_r['state_bytes']}, checkpoints {t['state_bytes']}")
    assert (got == got_r).all() and cig == cig_r
    _check_vs_oracle(got, cig, pairs, oracle)
    # per pair: records ~ score * W words, checkpoints ~ 2 sqrt(score * slots) * W
    assert t["state_bytes"] * 4 < t_r["state_bytes"]


def test_checkpoint_cigars_int32_offsets(ck_aligner, tsu, oracle):
    """Sequences over 32,000 bases (int32 ring cells in the snapshots)."""
    ds = tsu.Dataset.generate(2, seed=SEED + 7, len_min=40000, err_min=0.004)
    pairs = ds.pairs()
    say("checkpoint cigars int32 2x40000")
    ck_aligner.set_cigar_memory(ck_aligner.CIGAR_CHECKPOINTS, 5)
    got, cig, _ = ck_aligner.align_cigar(ck_aligner.batch(ds))
    _check_vs_oracle(got, cig, pairs, oracle)


def test_auto_mode_switches_under_budget(ck_aligner, tsu, oracle):
    """With a 64 MiB backtrace budget the automatic mode takes checkpoints
    for 10 kbp pairs at 10% (records would need ~70 MB per pair) and stays
    within the budget; with the default budget it keeps records."""
    pairs = tsu.Dataset.generate(6, seed=SEED + 29, len_min=10000, err_min=0.10).pairs()
    say("auto mode, budget 64 MiB")
    ck_aligner.set_cigar_memory(ck_aligner.CIGAR_AUTO, 0, 64 << 20)
    got, cig, t = ck_aligner.align_cigar(ck_aligner.batch(pairs))
    assert t["checkpoint_launches"] >= 1 and t["record_launches"] == 0
    assert t["state_bytes"] <= 64 << 20
    ck_aligner.set_cigar_memory(ck_aligner.CIGAR_AUTO)
    got_d, cig_d, t_d = ck_aligner.align_cigar(ck_aligner.batch(pairs))
    assert t_d["record_launches"] >= 1 and t_d["checkpoint_launches"] == 0
    assert (got == got_d).all() and cig == cig_d
    _check_vs_oracle(got, cig, pairs, oracle)


def test_other_penalties_checkpointed(tsu, oracle):
    """Penalties other than 4/6/2 (runtime ring depths): 1/1/1 and 3/5/1."""
    ds = tsu.Dataset.generate(60, seed=SEED + 31, len_min=200, len_max=2500, err_min=0.02, err_max=0.12)
    pairs = ds.pairs()
    for pen in [(1, 1, 1), (3, 5, 1)]:
        a = tsu.Aligner(*pen)
        try:
            say(f"checkpoint cigars penalties {pen}")
            a.set_cigar_memory(a.CIGAR_CHECKPOINTS, 4)
            got, cig, _ = a.align_cigar(a.batch(ds))
            _check_vs_oracle(got, cig, pairs, oracle, pen)
        finally:
            a.close()


def test_single_pass_step_bound_from_budget(ck_aligner, tsu, oracle):
    """A 150 MiB budget lowers the single pass's step bound below what the
    4096-diagonal window holds (1,064 of 2,046 scaled scores for 9 resident
    pairs): the 3-3.5 kbp pairs at 5% end within it, the unrelated pairs pass
    it and are done score first, with checkpoints at that budget.  Scores and
    CIGARs equal the oracle's."""
    rng = np.random.default_rng(43)
    pairs = tsu.Dataset.generate(6, seed=SEED + 41, len_min=3000, len_max=3500, err_min=0.05).pairs()
    for n in (2500, 3000, 3500):
        pairs.append(("".join(rng.choice(list("ACGT"), n)), "".join(rng.choice(list("ACGT"), n))))
    say("single pass under a 150 MiB budget, 9 pairs")
    ck_aligner.set_cigar_memory(ck_aligner.CIGAR_AUTO, 0, 150 << 20)
    try:
        got, cig, t = ck_aligner.align_cigar(ck_aligner.batch(pairs))
    finally:
        ck_aligner.set_cigar_memory(ck_aligner.CIGAR_AUTO)
    _check_vs_oracle(got, cig, pairs, oracle)
    assert t["record_launches"] >= 1 and t["checkpoint_launches"] >= 1
    assert t["state_bytes"] <= 150 << 20


def test_set_cigar_memory_rejects_bad_arguments(ck_aligner, tsu):
    for args in [(3, 0, 0), (-1, 0, 0), (0, -1, 0), (0, 0, -1)]:
        with pytest.raises(tsu.TsuError):
            ck_aligner.set_cigar_memory(*args)
    ck_aligner.set_cigar_memory()
