"""GPU: sequences longer than LDS holds.

Every kernel but one stages both sequences of a pair in LDS, which holds
about 327 kbp each.  Longer pairs run on the LDS kernel reading the packed
words in place from HBM (the reference reads its sequences from global
memory at any length).  Pairs of 327,329 bp (the first length past what LDS
holds), 400 kbp and 1 Mbp, and a ragged 350 / 351.5 kbp pair, at low
divergence so the oracle finishes in seconds, through the plain, batched,
CIGAR and served forms: scores and CIGARs equal the oracle's, and a batch
mixing them with short pairs (staged in LDS, on the register kernels) too.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
SEED = 0x4C4F4E47


def say(msg):
    print(f"\n[gpu] {msg}", flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def long_pairs(tsu):
    g = tsu.Dataset.generate
    pairs = (g(1, seed=SEED + 1, len_min=327_329, err_min=0.002).pairs() +
             g(2, seed=SEED + 2, len_min=400_000, err_min=0.001, err_max=0.003).pairs() +
             g(1, seed=SEED + 3, len_min=1_000_000, err_min=0.0005).pairs())
    p, t = g(1, seed=SEED + 4, len_min=350_000, err_min=0.001).pairs()[0]
    pairs.append((p, t + t[:1500]))  # a 1,500-base gap at the end
    return pairs


@pytest.fixture(scope="module")
def want(oracle, long_pairs):
    return oracle.wfa_batch(long_pairs, 4, 6, 2, cigar=True)


def test_plain_reads_sequences_from_hbm(aligner, long_pairs, want):
    sc, _, _ = want
    assert min(max(len(p), len(t)) for p, t in long_pairs) > 327_328
    W = 2 * ((int(sc.max()) - 6) // 2) + 3
    say(f"plain, {len(long_pairs)} long pairs, W={W}")
    got, _ = aligner.align_plain(aligner.batch(long_pairs), W)
    assert (got == sc).all(), (got, sc)


def test_batched_and_cigar(aligner, tsu, long_pairs, want, oracle):
    sc, ops, _ = want
    short = tsu.Dataset.generate(300, seed=SEED + 5, len_min=100, len_max=20_000, err_min=0.01, err_max=0.1).pairs()
    pairs = long_pairs + short
    ws, wops, _ = oracle.wfa_batch(short, 4, 6, 2, cigar=True)
    b = aligner.batch(pairs)
    say(f"batched, {len(long_pairs)} long + {len(short)} short pairs")
    got, _ = aligner.align_batched(b)
    assert (got[:len(long_pairs)] == sc).all()
    assert (got[len(long_pairs):] == ws).all()
    say("CIGARs")
    got2, cig, t = aligner.align_cigar(b)
    assert (got2 == got).all()
    allops = list(ops) + list(wops)
    bad = [i for i in range(len(pairs)) if cig[i] != oracle.rle(allops[i])]
    assert not bad, bad[:10]
    assert "lds" in t["kernels"] or "lds_ck" in t["kernels"]


def test_checkpointed_cigar(tsu, long_pairs, want, oracle):
    sc, ops, _ = want
    a = tsu.Aligner(4, 6, 2)
    try:
        a.set_cigar_memory(a.CIGAR_CHECKPOINTS, 64)
        say("checkpointed CIGARs of the long pairs")
        got, cig, t = a.align_cigar(a.batch(long_pairs))
        assert t["kernels"] == {"lds_ck"}, t["kernels"]
        assert (got == sc).all()
        assert all(cig[i] == oracle.rle(ops[i]) for i in range(len(long_pairs)))
    finally:
        a.close()


def test_served(aligner, tsu, long_pairs, want):
    sc, _, _ = want
    srv = tsu.Server(aligner, hbm_budget_bytes=4 << 30, max_inflight=2, max_batch_pairs=64)
    say("served long pairs")
    tickets = [srv.submit([pr]) for pr in long_pairs]
    for i, tk in enumerate(tickets):
        st, got, _ = srv.wait(tk, 1)
        assert st == tsu.TSU_OK and int(got[0]) == int(sc[i]), i
    assert srv.close() == tsu.TSU_OK
