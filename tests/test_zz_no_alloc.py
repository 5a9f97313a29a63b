"""tsu_align_batched_noalloc against tsu_align_batched, through the C-ABI.

The no-allocation entry point keeps its work counters and queue order in a
file-scope device array and takes the caller's device arrays; its scores
must equal the batched form's bit for bit, on two calls in a row (the
second smaller, so stale scratch contents from the first would show), and
it refuses what its scratch cannot hold.
"""
import ctypes
import random
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
SEED = 0x5453554E


def say(msg):
    print(f"\n[gpu] {msg}", flush=True)
    sys.stdout.flush()


def upload(tsu, a, words, pairs):
    """A batch made from caller-packed arrays, and its device arrays."""
    lib = tsu.library()
    h = ctypes.c_void_p()
    P = ctypes.POINTER
    r = lib.tsu_batch_create(a._h, words.ctypes.data_as(P(ctypes.c_uint32)), len(words),
                             pairs.ctypes.data_as(P(tsu.Pair)), len(pairs), ctypes.byref(h))
    assert r == tsu.TSU_OK, tsu.strerror(r)
    b = tsu.Batch(a, h)
    dw, nw, dp, dsc = ctypes.c_void_p(), ctypes.c_int64(), ctypes.c_void_p(), ctypes.c_void_p()
    assert lib.tsu_batch_device(h, ctypes.byref(dw), ctypes.byref(nw), ctypes.byref(dp), ctypes.byref(dsc)) == 0
    assert nw.value == len(words)
    return b, dw, dp, dsc


def noalloc(tsu, a, dw, n_words, dp, pairs, dsc, n=None):
    lib = tsu.library()
    n = len(pairs) if n is None else n
    scores = np.full(max(len(pairs), 1), -99, np.int32)
    t = tsu.Timing()
    r = lib.tsu_align_batched_noalloc(a._h, dw, n_words, dp, pairs.ctypes.data_as(ctypes.POINTER(tsu.Pair)), n, dsc,
                                      scores.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), ctypes.byref(t))
    return r, scores[:len(pairs)], t


def test_noalloc_equals_batched_twice(tsu):
    a = tsu.Aligner(4, 6, 2)
    try:
        big = tsu.Dataset.generate(3000, seed=SEED + 90, len_min=100, len_max=5000, err_min=0.01, err_max=0.15)
        small = tsu.Dataset.generate(400, seed=SEED + 91, len_min=100, len_max=1200, err_min=0.02, err_max=0.10)
        for name, ds in (("3000 pairs of 100-5000 bp", big), ("400 pairs of 100-1200 bp", small)):
            words, pairs = ds.packed()
            b_in, dw, dp, _ = upload(tsu, a, words, pairs)
            b_out, _, _, dsc = upload(tsu, a, words, pairs)   # its score slots receive the output
            want, _ = a.align_batched(b_in)
            say(f"no-allocation batched form, {name}")
            r, got, t = noalloc(tsu, a, dw, len(words), dp, pairs, dsc)
            assert r == tsu.TSU_OK, a._lib.tsu_last_error(a._h)
            assert (got == want).all(), np.nonzero(got != want)[0][:10]
            assert t.launches >= 1 and t.kernel_ms > 0 and t.pairs == len(pairs)
            b_in.close()
            b_out.close()
    finally:
        a.close()


def test_noalloc_limits(tsu):
    a = tsu.Aligner(4, 6, 2)
    try:
        rng = random.Random(SEED)
        wide = ("".join(rng.choice("ACGT") for _ in range(20000)), "".join(rng.choice("ACGT") for _ in range(20000)))
        ds = tsu.Dataset.from_pairs([("ACGTACGT", "ACGTTCGT"), wide, ("GATTACA", "GATACA")])
        words, pairs = ds.packed()
        b, dw, dp, dsc = upload(tsu, a, words, pairs)
        lim = tsu.library().tsu_align_batched_noalloc
        say("no-allocation batched form: more pairs than the scratch holds")
        r, _, _ = noalloc(tsu, a, dw, len(words), dp, pairs, dsc, n=1048577)
        assert r == tsu.TSU_EINVAL
        assert b"at most 1048576" in a._lib.tsu_last_error(a._h)
        say("no-allocation batched form: a pair wider than the register windows")
        r, got, _ = noalloc(tsu, a, dw, len(words), dp, pairs, dsc)
        assert r == tsu.TSU_EKOVERFLOW
        assert got[1] == tsu.TSU_EKOVERFLOW
        assert got[0] == 4 and got[2] == 8
        assert lim(a._h, None, 0, None, None, 0, None, None, None) == tsu.TSU_OK
        b.close()
    finally:
        a.close()
