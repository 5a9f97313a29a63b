"""GPU: the continuous-batching server (tsu_serve_*).

Requests arrive from several threads while others are on the device; every
request's scores must equal the oracle's and the same request's scores when
it is served alone (no cross-request state).
"""
import threading

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
SEED = 0x5453554E + 5


def say(msg):
    print(f"\n[gpu] {msg}", flush=True)


def make_requests(tsu, n_req, max_pairs=64, len_max=3000, seed=SEED):
    rng = np.random.default_rng(seed)
    reqs = []
    for r in range(n_req):
        k = int(rng.integers(1, max_pairs + 1))
        ds = tsu.Dataset.generate(k, seed=seed * 1000 + r, len_min=100, len_max=len_max, err_min=0.01, err_max=0.15)
        reqs.append(ds.pairs())
    return reqs


def test_serve_mix_equals_alone_and_oracle(aligner, tsu, oracle):
    reqs = make_requests(tsu, 40)
    say(f"serve mix: {len(reqs)} requests, {sum(map(len, reqs))} pairs")
    srv = tsu.Server(aligner, hbm_budget_bytes=1 << 30, max_inflight=4, max_batch_pairs=256)
    results = {}
    errors = []

    def client(idx):
        try:
            for i in idx:
                t = srv.submit(reqs[i])
                st, sc, stats = srv.wait(t, len(reqs[i]))
                results[i] = (st, sc.copy(), stats)
        except Exception as ex:  # surfaced below
            errors.append(ex)

    threads = [threading.Thread(target=client, args=(list(range(j, len(reqs), 4)),)) for j in range(4)]
    for th in threads:
        th.start()
    for th in threads:
        th.join()
    assert not errors, errors
    for i, pairs in enumerate(reqs):
        st, sc, stats = results[i]
        assert st == tsu.TSU_OK
        want, _, _ = oracle.wfa_batch(pairs, 4, 6, 2)
        assert (sc == want).all(), i
        assert stats["submit_ns"] <= stats["admit_ns"] <= stats["done_ns"]
    say("serve alone")
    for i in (0, 7, 23):
        t = srv.submit(reqs[i])
        st, sc, _ = srv.wait(t, len(reqs[i]))
        assert st == tsu.TSU_OK and (sc == results[i][1]).all()
    assert srv.close() == tsu.TSU_OK


def test_serve_poll_budget_and_empty(aligner, tsu):
    srv = tsu.Server(aligner, hbm_budget_bytes=2 << 20, max_inflight=2, max_batch_pairs=64)
    say("serve poll / empty / budget")
    t = srv.submit([])
    st, sc, stats = srv.wait(t, 0)
    assert st == tsu.TSU_OK and stats["n_pairs"] == 0
    pairs = tsu.Dataset.generate(8, seed=3, len_min=200, err_min=0.02).pairs()
    t = srv.submit(pairs)
    while True:
        st, sc, _ = srv.poll(t, len(pairs))
        if st != tsu.TSU_EAGAIN:
            break
    assert st == tsu.TSU_OK and len(sc) == 8
    # 600 pairs of 2 x 8 kbp need 2.4 MB packed: more than the whole budget
    big = tsu.Dataset.generate(600, seed=4, len_min=8000, err_min=0.0).pairs()
    with pytest.raises(tsu.TsuError) as ei:
        srv.submit(big)
    assert ei.value.status == tsu.TSU_EBUDGET
    # a pair of 400 kbp: its flight reserves the ring of every widening up to
    # the full 800,003 diagonals (32 MB), more than the whole budget
    huge = tsu.Dataset.generate(1, seed=4, len_min=400_000, err_min=0.0).pairs()
    with pytest.raises(tsu.TsuError) as ei:
        srv.submit(huge)
    assert ei.value.status == tsu.TSU_EBUDGET
    assert srv.close() == tsu.TSU_OK


def test_serve_overflow_pairs_are_widened(aligner, tsu, oracle):
    # unrelated random sequences need the full diagonal range: the first
    # capacity overflows and the flight re-launches them wider
    rng = np.random.default_rng(9)
    pairs = [("".join(rng.choice(list("ACGT"), 700)), "".join(rng.choice(list("ACGT"), 650))) for _ in range(6)]
    say("serve overflow widening")
    srv = tsu.Server(aligner, hbm_budget_bytes=256 << 20, max_inflight=2, max_batch_pairs=64)
    t = srv.submit(pairs)
    st, sc, _ = srv.wait(t, len(pairs))
    want, _, _ = oracle.wfa_batch(pairs, 4, 6, 2)
    assert st == tsu.TSU_OK and (sc == want).all()
    assert srv.close() == tsu.TSU_OK


def test_serve_submit_packed(aligner, tsu, oracle):
    """Pre-packed requests give the same scores; pairs outside the given
    words are refused."""
    ds = tsu.Dataset.generate(40, seed=21, len_min=100, len_max=3000, err_min=0.01, err_max=0.15)
    words, pairs = ds.packed()
    want, _, _ = oracle.wfa_batch(ds.pairs(), 4, 6, 2)
    say("serve packed")
    srv = tsu.Server(aligner, hbm_budget_bytes=256 << 20, max_inflight=2, max_batch_pairs=64)
    t1 = srv.submit_packed(words, pairs[:25])
    t2 = srv.submit_packed(words, pairs[25:])
    st1, sc1, _ = srv.wait(t1, 25)
    st2, sc2, _ = srv.wait(t2, 15)
    assert st1 == tsu.TSU_OK and st2 == tsu.TSU_OK
    assert (np.concatenate([sc1, sc2]) == want).all()
    with pytest.raises(tsu.TsuError) as ei:
        srv.submit_packed(words[:10], pairs[30:31])
    assert ei.value.status == tsu.TSU_EINVAL
    assert srv.close() == tsu.TSU_OK
