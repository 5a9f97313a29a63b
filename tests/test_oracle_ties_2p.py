"""Pins the 2-piece oracle's choice among co-optimal CIGARs.

The walk (orc_wfa2p_cigar) follows WFA2-lib's affine-2p backtrace with the
nine backtrace types of common.h:67-77.  Pinned here:
  - every CIGAR re-scores to the optimum of the 2-piece Gotoh DP, on the
    hand-derived cases and on random pairs under several penalty sets;
  - with the second piece absent (its gap never cheaper than the score),
    with the first absent, and with two equal pieces, the 2-piece walk gives
    the 1-piece walk's CIGAR, whose choice is pinned by the hand-derived
    1-piece cases (tie_cases.TIES_462 / TIES_531);
  - one hand-derived case per adjacent pair of the nine-type order
    (tie_cases.TIES_2P): the true order gives the expected CIGAR, swapping
    that pair gives the other choice's, and no other adjacent swap changes
    it; so every order other than the reference's (same relative order of
    all nine types) changes at least one case, which is also checked on a
    sample of random orders.
"""
import random

import pytest

from tie_cases import ALL_TIES, TIES_2P

KINDS2 = ("M", "D2_ext", "D2_open", "D1_ext", "D1_open", "I2_ext", "I2_open", "I1_ext", "I1_open")
BIG = 10 ** 6  # an open penalty no test score reaches: that piece never appears


def test_priority_order_is_the_references(oracle):
    # bt_M 9, bt_D2_ext 8, bt_D2_open 7, bt_D1_ext 6, bt_D1_open 5, bt_I2_ext 4,
    # bt_I2_open 3, bt_I1_ext 2, bt_I1_open 1 (common.h:67-77)
    assert oracle.PRIORITY_2P_KINDS == KINDS2
    assert oracle.PRIORITY_2P == (9, 8, 7, 6, 5, 4, 3, 2, 1)


def swapped(i, j):
    pr = list((9, 8, 7, 6, 5, 4, 3, 2, 1))
    pr[i], pr[j] = pr[j], pr[i]
    return pr


@pytest.mark.parametrize("case", TIES_2P, ids=lambda c: f"{c[4][0]}-{c[4][1]}")
def test_hand_derived_2p_tie_cigars(oracle, case):
    p, t, pen, want, (a, b), other = case
    s, ops = oracle.wfa2p_cigar(p, t, *pen)
    assert oracle.rle(ops) == want
    best = oracle.gotoh2p(p, t, *pen)
    assert s == best == oracle.wfa2p(p, t, *pen)[0]
    assert oracle.rescore2p(oracle.expand(want), p, t, *pen) == best
    assert oracle.rescore2p(oracle.expand(other), p, t, *pen) == best
    assert other != want


@pytest.mark.parametrize("case", TIES_2P, ids=lambda c: f"{c[4][0]}-{c[4][1]}")
def test_each_2p_case_is_decided_by_its_adjacent_pair(oracle, case):
    p, t, pen, want, (a, b), other = case
    ia, ib = KINDS2.index(a), KINDS2.index(b)
    assert ib == ia + 1  # adjacent in the order
    for i in range(8):
        _, ops = oracle.wfa2p_cigar(p, t, *pen, prio=swapped(i, i + 1))
        got = oracle.rle(ops)
        if i == ia:
            assert got == other
        else:
            assert got == want, (KINDS2[i], KINDS2[i + 1], got)


def test_every_adjacent_pair_has_a_case():
    pairs = {(c[4][0], c[4][1]) for c in TIES_2P}
    assert pairs == {(KINDS2[i], KINDS2[i + 1]) for i in range(8)}


def test_other_orders_change_some_case(oracle):
    rng = random.Random(7)
    vals = list(range(1, 16))
    for _ in range(300):
        pr = rng.sample(vals, 9)
        keeps = all(pr[i] > pr[i + 1] for i in range(8))
        changed = False
        for p, t, pen, want, _, _ in TIES_2P:
            _, ops = oracle.wfa2p_cigar(p, t, *pen, prio=pr)
            if oracle.rle(ops) != want:
                changed = True
                break
        assert changed != keeps, pr


@pytest.mark.parametrize("case", ALL_TIES, ids=lambda c: f"{c[0]}-{c[1]}-{c[4][0]}-{c[4][1]}")
def test_2p_walk_reduces_to_the_1p_walk(oracle, case):
    p, t, (x, o, e), want, _, _ = case
    for pen in ((x, o, e, BIG, 1), (x, BIG, 1, o, e), (x, o, e, o, e)):
        s, ops = oracle.wfa2p_cigar(p, t, *pen)
        assert oracle.rle(ops) == want, pen
        assert s == oracle.gotoh(p, t, x, o, e)


def test_2p_cigars_rescore_to_gotoh(oracle):
    rng = random.Random(11)
    pens = [(4, 6, 2, 24, 1), (2, 4, 2, 10, 1), (3, 2, 3, 8, 1), (4, 1, 3, 7, 1), (8, 5, 4, 4, 5)]
    for _ in range(400):
        p = "".join(rng.choice("ACGT") for _ in range(rng.randint(0, 24)))
        t = "".join(rng.choice("ACGT") for _ in range(rng.randint(0, 24)))
        pen = rng.choice(pens)
        s, ops = oracle.wfa2p_cigar(p, t, *pen)
        assert s == oracle.gotoh2p(p, t, *pen)
        assert oracle.rescore2p(ops, p, t, *pen) == s
    for _ in range(300):  # reduces to the 1-piece walk on random pairs too
        p = "".join(rng.choice("ACG") for _ in range(rng.randint(0, 14)))
        t = "".join(rng.choice("ACG") for _ in range(rng.randint(0, 14)))
        x, o, e = rng.choice([(4, 6, 2), (5, 3, 1), (2, 3, 1)])
        s1, c1, _ = oracle.wfa(p, t, x, o, e, cigar=True)
        s2, c2 = oracle.wfa2p_cigar(p, t, x, o, e, o, e)
        assert (s1, c1) == (s2, c2)
