"""GPU parity for gap-affine 2-piece penalties (a gap of length L costs
min(o1 + L e1, o2 + L e2)): libtsunami_hip.so against the 2-piece oracle
(oracle/wfa2p_oracle.c, itself pinned by a 2-piece Gotoh DP), bit-exact."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
SEED = 0x5453554E
# extensions (e1, e2) after the gcd: (2, 1), (1, 1), (3, 1), (2, 1), (2, 1)
# take the single-wave register kernel in windows of up to 1,024 diagonals;
# (3, 2) the LDS-ring kernel
PEN_2P = [(4, 6, 2, 24, 1), (4, 6, 2, 12, 2), (5, 1, 3, 10, 1), (3, 4, 2, 20, 1), (6, 5, 2, 30, 1),
          (4, 5, 3, 20, 2)]


def register_kernel(pen):
    g = math.gcd(*pen)
    return (pen[2] // g, pen[4] // g) in ((1, 1), (2, 1), (3, 1))


def say(msg):
    print(f"\n[gpu] {msg}", flush=True)


def pairs_for(tsu, seed):
    ds = tsu.Dataset.generate(500, seed=seed, len_min=20, len_max=400, err_min=0.0, err_max=0.2)
    pairs = ds.pairs()
    rng = np.random.default_rng(seed)
    for _ in range(20):  # unrelated random pairs: long gaps, the second piece wins
        pairs.append(("".join(rng.choice(list("ACGT"), int(rng.integers(1, 300)))),
                      "".join(rng.choice(list("ACGT"), int(rng.integers(1, 300))))))
    pairs += [("", ""), ("A", ""), ("", "ACGT" * 10), ("A" * 200, "C" * 30), ("ACGT" * 50, "ACGT" * 50),
              ("ACGT" * 30 + "T" * 60 + "ACGT" * 30, "ACGT" * 60)]
    return pairs


@pytest.mark.parametrize("pen", PEN_2P)
def test_two_piece_plain_and_batched(tsu, oracle, pen):
    """Plain (capacity 1000 covers every pair here) and batched forms equal the
    oracle; a narrow capacity reports overflow for the wide pairs only."""
    a = tsu.Aligner(pen[0], pen[1], pen[2], gap_opening2=pen[3], gap_extension2=pen[4])
    try:
        pairs = pairs_for(tsu, SEED + 20 + sum(pen))
        want, _ = oracle.wfa2p_batch(pairs, *pen)
        b = a.batch(pairs)
        say(f"2-piece {pen} plain W=1000")
        got, _ = a.align_plain(b, 1000)
        assert (got == want).all(), pen
        say(f"2-piece {pen} batched")
        got, _ = a.align_batched(b)
        assert (got == want).all(), pen
        say(f"2-piece {pen} plain W=40")
        got, _ = a.align_plain(b, 40, raise_on_error=False)
        ok = got != tsu.TSU_EKOVERFLOW
        assert ok.sum() > 0 and (~ok).sum() > 0
        assert (got[ok] == want[ok]).all(), pen
    finally:
        a.close()


@pytest.mark.parametrize("pen", PEN_2P)
def test_two_piece_cigars(tsu, oracle, pen):
    """2-piece CIGARs (records on the 2-piece kernel, walked with the nine
    backtrace types) equal the 2-piece oracle's, itself pinned by
    tie_cases.TIES_2P and Gotoh re-scoring."""
    a = tsu.Aligner(pen[0], pen[1], pen[2], gap_opening2=pen[3], gap_extension2=pen[4])
    try:
        pairs = pairs_for(tsu, SEED + 40 + sum(pen))
        say(f"2-piece {pen} CIGARs, {len(pairs)} pairs")
        got, cig, t = a.align_cigar(a.batch(pairs))
        if register_kernel(pen):  # the widest pairs need the LDS-ring kernel
            assert "2p_reg" in t["kernels"] and t["kernels"] <= {"2p", "2p_reg"}, t["kernels"]
        else:
            assert t["kernels"] == {"2p"}, t["kernels"]
        for i, (p, q) in enumerate(pairs):
            s, ops = oracle.wfa2p_cigar(p, q, *pen)
            assert got[i] == s, i
            assert cig[i] == oracle.rle(ops), (i, p, q)
    finally:
        a.close()


@pytest.mark.parametrize("pen", [p for p in PEN_2P if register_kernel(p)])
def test_two_piece_register_kernel_1000bp(tsu, oracle, pen):
    """Pairs of the benchmark's shape (1000 bp, 1-8% mixed edits) in windows
    of 256, 512, 768 and 1,024 diagonals of the register kernel: plain scores
    and CIGARs equal the oracle's (and a window too narrow for a pair
    reports TSU_EKOVERFLOW for it, as the LDS-ring kernel does)."""
    a = tsu.Aligner(pen[0], pen[1], pen[2], gap_opening2=pen[3], gap_extension2=pen[4])
    try:
        pairs = tsu.Dataset.generate(300, seed=SEED + 80 + sum(pen), len_min=1000, err_min=0.01,
                                     err_max=0.08).pairs()
        want, _ = oracle.wfa2p_batch(pairs, *pen)
        b = a.batch(pairs)
        for W in (256, 512, 768, 1024):
            say(f"2-piece {pen} register kernel, plain W={W}")
            got, _ = a.align_plain(b, W, raise_on_error=False)
            ok = got != tsu.TSU_EKOVERFLOW
            assert (got[ok] == want[ok]).all(), (pen, W)
            if W == 1024:
                assert ok.sum() > len(pairs) // 2
        say(f"2-piece {pen} register kernel, CIGARs")
        got, cig, t = a.align_cigar(b)
        assert "2p_reg" in t["kernels"], t["kernels"]
        assert (got == want).all()
        for i in range(0, len(pairs), 10):
            s, ops = oracle.wfa2p_cigar(pairs[i][0], pairs[i][1], *pen)
            assert cig[i] == oracle.rle(ops), i
    finally:
        a.close()


def test_two_piece_tie_cigars(tsu, oracle):
    """The hand-derived 2-piece tie cases, alone and behind a common prefix
    (which only prefixes matches: every cell of the longer pair is the
    shorter pair's, shifted along the prefix)."""
    from tie_cases import TIES_2P
    from test_gpu_cigar_ties import merge_prefix, prefixed
    for pen in sorted({c[2] for c in TIES_2P}):
        cases = [c for c in TIES_2P if c[2] == pen]
        a = tsu.Aligner(pen[0], pen[1], pen[2], gap_opening2=pen[3], gap_extension2=pen[4])
        try:
            for n in (0, 300):
                pairs, want = prefixed(cases, n, 5 + n)
                say(f"2-piece ties {pen}, prefix {n}: {len(pairs)} pairs")
                got, cig, t = a.align_cigar(a.batch(pairs))
                assert cig == want, (pen, n)
        finally:
            a.close()


def test_two_piece_served(tsu, oracle):
    """The served form with 2-piece penalties: requests submitted from two
    threads while others are on the device equal the oracle and the same
    request served alone."""
    import threading
    pen = (4, 6, 2, 24, 1)
    a = tsu.Aligner(pen[0], pen[1], pen[2], gap_opening2=pen[3], gap_extension2=pen[4])
    try:
        rng = np.random.default_rng(SEED + 60)
        reqs = []
        for r in range(24):
            k = int(rng.integers(1, 65))
            ds = tsu.Dataset.generate(k, seed=SEED + 600 + r, len_min=50, len_max=1000, err_min=0.01, err_max=0.15)
            reqs.append(ds.pairs())
        srv = tsu.Server(a, hbm_budget_bytes=1 << 30, max_inflight=4, max_batch_pairs=128)
        results, errors = {}, []

        def client(idx):
            try:
                for i in idx:
                    tk = srv.submit(reqs[i])
                    results[i] = srv.wait(tk, len(reqs[i]))
            except Exception as ex:  # surfaced below
                errors.append(ex)

        say(f"2-piece served: {len(reqs)} requests, {sum(map(len, reqs))} pairs")
        ths = [threading.Thread(target=client, args=(list(range(j, len(reqs), 2)),)) for j in range(2)]
        for th in ths:
            th.start()
        for th in ths:
            th.join()
        assert not errors, errors
        for i, pairs in enumerate(reqs):
            st, sc, _ = results[i]
            assert st == tsu.TSU_OK
            want, _ = oracle.wfa2p_batch(pairs, *pen)
            assert (sc == want).all(), i
        say("2-piece served alone")
        for i in (0, 5, 17):
            tk = srv.submit(reqs[i])
            st, sc, _ = srv.wait(tk, len(reqs[i]))
            assert st == tsu.TSU_OK and (sc == results[i][1]).all()
        assert srv.close() == tsu.TSU_OK
    finally:
        a.close()


def test_two_piece_beyond_lds_reports_overflow(tsu):
    """A 2-piece pair whose rings at the capacity it needs exceed the LDS (the
    2-piece kernel has no HBM ring): the batched form widens up to the widest
    window that fits and then fails with TSU_EKOVERFLOW."""
    rng = np.random.default_rng(SEED + 70)
    wide = ("".join(rng.choice(list("ACGT"), 6000)), "".join(rng.choice(list("ACGT"), 6000)))
    a = tsu.Aligner(4, 6, 2, gap_opening2=24, gap_extension2=1)
    try:
        b = a.batch([("ACGTACGT", "ACGAACGT"), wide])
        say("2-piece pair beyond the LDS")
        with pytest.raises(tsu.TsuError) as ei:
            a.align_batched(b)
        assert ei.value.status == tsu.TSU_EKOVERFLOW
    finally:
        a.close()
