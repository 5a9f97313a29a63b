"""GPU: the mixed-length and served workloads end to end.

cfg4 (500,000 pairs, length log-uniform in [100, 20,000] bp, error uniform
in [1%, 15%], 4/6/2): a 3,000-pair draw of the same generator through the
autotuner, the batched form and the CIGAR form against the oracle; the
full 500,000 pairs through the batched form with size-independent checks
(every score non-negative, even, within the bound of min(p,t) mismatches
plus one gap, identical on a re-run) and a 2,000-pair oracle sample.

cfg5 (Poisson arrivals of requests of U{1..256} pairs, 100-10,000 bp,
1-15%): 60 requests submitted open-loop at exponential gaps, so several
are in flight at once; each must equal the oracle and equal the same
request served alone.
"""
import sys
import threading
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
SEED = 0x5453554E


def say(msg):
    print(f"\n[gpu] {msg}", flush=True)
    sys.stdout.flush()


def cfg4(tsu, n):
    return tsu.Dataset.generate(n, seed=SEED + 4, len_min=100, len_max=20_000, err_min=0.01, err_max=0.15)


def test_cfg4_sample_autotuned_batched_and_cigars(tsu, oracle):
    ds = cfg4(tsu, 3000)
    pairs = ds.pairs()
    assert max(max(len(p), len(t)) for p, t in pairs) > 16_000   # the long tail is in the draw
    want, ops, _ = oracle.wfa_batch(pairs, 4, 6, 2, cigar=True)
    a = tsu.Aligner(4, 6, 2)
    try:
        b = a.batch(ds)
        say("cfg4 sample: autotune 3000 pairs")
        table = a.autotune(b, 512)
        assert table
        say("cfg4 sample: batched")
        got, t = a.align_batched(b)
        assert (got == want).all(), np.nonzero(got != want)[0][:10]
        say("cfg4 sample: CIGARs")
        got2, cig, t2 = a.align_cigar(b)
        assert (got2 == want).all()
        bad = [i for i in range(len(pairs)) if cig[i] != oracle.rle(ops[i])]
        assert not bad, bad[:10]
        say(f"cfg4 sample: CIGAR kernels {sorted(t2['kernels'])}")
    finally:
        a.close()


def test_cfg4_full_size_properties(tsu, oracle):
    ds = cfg4(tsu, 500_000)
    pl, tl = ds.lengths()
    a = tsu.Aligner(4, 6, 2)
    try:
        b = a.batch(ds)
        say("cfg4 full size: autotune + batched, 500000 pairs")
        a.autotune(b, 2048)
        got, t = a.align_batched(b)
        assert (got >= 0).all()
        assert (got % 2 == 0).all()                      # every penalty is even
        bound = 4 * np.minimum(pl, tl) + np.where(pl != tl, 6 + 2 * np.abs(pl.astype(np.int64) - tl), 0)
        assert (got <= bound).all()
        say("cfg4 full size: re-run")
        again, _ = a.align_batched(b)
        assert (again == got).all()
        want, _, _ = oracle.wfa_batch([ds.pair(i) for i in range(2000)], 4, 6, 2)
        assert (got[:2000] == want).all()
    finally:
        a.close()


def test_cfg5_poisson_requests_equal_alone_and_oracle(tsu, oracle):
    rng = np.random.default_rng(SEED + 5)
    sizes = rng.integers(1, 257, 60)
    ds = tsu.Dataset.generate(int(sizes.sum()), seed=SEED + 55, len_min=100, len_max=10_000,
                              err_min=0.01, err_max=0.15)
    firsts = np.concatenate([[0], np.cumsum(sizes)[:-1]])
    reqs = [[ds.pair(int(f) + j) for j in range(int(k))] for f, k in zip(firsts, sizes)]
    want, _, _ = oracle.wfa_batch(ds.pairs(), 4, 6, 2)
    a = tsu.Aligner(4, 6, 2)
    try:
        srv = tsu.Server(a, hbm_budget_bytes=4 << 30, max_inflight=4, max_batch_pairs=2048)
        gaps = rng.exponential(0.004, len(reqs))          # about 250 requests/s, open loop
        tickets = [None] * len(reqs)
        ready = [threading.Event() for _ in reqs]
        errors = []

        def submitter():
            try:
                t0 = time.perf_counter()
                due = 0.0
                for i, r in enumerate(reqs):
                    due += gaps[i]
                    d = t0 + due - time.perf_counter()
                    if d > 0:
                        time.sleep(d)
                    tickets[i] = srv.submit(r)
                    ready[i].set()
            except Exception as ex:  # surfaced below
                errors.append(ex)
                for ev in ready:
                    ev.set()

        say(f"cfg5: {len(reqs)} Poisson requests, {int(sizes.sum())} pairs")
        th = threading.Thread(target=submitter)
        th.start()
        mixed, overlap = [], 0
        for i, r in enumerate(reqs):
            ready[i].wait()
            assert not errors, errors
            st, sc, stats = srv.wait(tickets[i], len(r))
            assert st == tsu.TSU_OK, (i, st)
            mixed.append((sc.copy(), stats))
        th.join()
        # requests whose time on the server overlapped another's
        spans = [(s["submit_ns"], s["done_ns"]) for _, s in mixed]
        for i, (s0, e0) in enumerate(spans):
            if any(j != i and s1 < e0 and s0 < e1 for j, (s1, e1) in enumerate(spans)):
                overlap += 1
        say(f"cfg5: {overlap} of {len(reqs)} requests overlapped another on the server")
        assert overlap >= len(reqs) // 2
        for i, r in enumerate(reqs):
            f = int(firsts[i])
            assert (mixed[i][0] == want[f:f + len(r)]).all(), i
        say("cfg5: every request alone")
        for i, r in enumerate(reqs):
            t = srv.submit(r)
            st, sc, _ = srv.wait(t, len(r))
            assert st == tsu.TSU_OK and (sc == mixed[i][0]).all(), i
        assert srv.close() == tsu.TSU_OK
    finally:
        a.close()
