"""Pins the oracle's choice among co-optimal CIGARs.

The cases of tie_cases.py carry CIGARs derived by hand from WFA2-lib's
published backtrace (wavefront_backtrace_affine) and the priority order the
reference restates in common.h:67-77.  Here:
  - the oracle reproduces every hand-derived CIGAR, and both the expected
    CIGAR and the one the other choice gives re-score to the optimum
    (Gotoh and exhaustive enumeration agree on it);
  - each case is decided by its one comparison: exactly the priority orders
    that invert that comparison change its CIGAR, and swapping the two
    priorities gives the other choice's CIGAR;
  - every one of the 119 other orders of the five source types changes at
    least one case, under each of the two penalty sets, and so do the three
    orders of the review (I before D, M last, open before extend);
  - gaps in homopolymer runs go to the right end of the run (largest
    offset first);
  - prefixing both sequences with the same bases only prefixes the CIGAR
    with matches (what the GPU tests rely on to reach each kernel).
"""
import itertools
import random

import pytest

from tie_cases import ALL_TIES, HOMOPOLYMER_462, TIES_462, TIES_531
from test_oracle import brute

KINDS = ("M", "D_ext", "D_open", "I_ext", "I_open")


def all_orders(oracle):
    """Every assignment of the five priority values to the five kinds."""
    vals = sorted(oracle.PRIORITY, reverse=True)
    out = []
    for order in itertools.permutations(range(5)):
        pr = [0] * 5
        for rank, kind in enumerate(order):
            pr[kind] = vals[rank]
        out.append(tuple(pr))
    return out


def rle(oracle, p, t, pen, prio):
    s, ops = oracle.wfa_prio(p, t, *pen, prio)
    return s, oracle.rle(ops)


def test_priority_order_is_the_references(oracle):
    # bt_M 9, bt_D1_ext 6, bt_D1_open 5, bt_I1_ext 2, bt_I1_open 1 (common.h:67-77)
    assert oracle.PRIORITY_KINDS == KINDS
    assert oracle.PRIORITY == (9, 6, 5, 2, 1)


@pytest.mark.parametrize("case", ALL_TIES, ids=lambda c: f"{c[0]}-{c[1]}-{c[4][0]}-{c[4][1]}")
def test_hand_derived_tie_cigars(oracle, case):
    p, t, pen, want, (a, b), other = case
    s, ops, _ = oracle.wfa(p, t, *pen, cigar=True)
    assert oracle.rle(ops) == want
    best = oracle.gotoh(p, t, *pen)
    assert s == best == brute(p, t, *pen)
    assert oracle.rescore(oracle.expand(want), p, t, *pen) == best
    assert oracle.rescore(oracle.expand(other), p, t, *pen) == best
    assert other != want
    # the default-priority entry point and the parameterised one agree
    assert rle(oracle, p, t, pen, oracle.PRIORITY) == (s, want)
    # the nulled variant (neighbour-synchronised kernels) walks to the same CIGAR
    assert oracle.wfa_nulled(p, t, *pen, cigar=True) == (s, ops)


@pytest.mark.parametrize("case", ALL_TIES, ids=lambda c: f"{c[0]}-{c[1]}-{c[4][0]}-{c[4][1]}")
def test_each_case_is_decided_by_one_comparison(oracle, case):
    p, t, pen, want, (a, b), other = case
    ia, ib = KINDS.index(a), KINDS.index(b)
    default = oracle.PRIORITY
    for pr in all_orders(oracle):
        inverted = (pr[ia] > pr[ib]) != (default[ia] > default[ib])
        s, got = rle(oracle, p, t, pen, pr)
        assert (got != want) == inverted, (pr, got)
    # the order with only those two swapped gives the other choice's CIGAR
    swapped = list(default)
    swapped[ia], swapped[ib] = swapped[ib], swapped[ia]
    assert rle(oracle, p, t, pen, swapped)[1] == other


@pytest.mark.parametrize("cases", [TIES_462, TIES_531], ids=["4/6/2", "5/3/1"])
def test_every_other_priority_order_changes_a_cigar(oracle, cases):
    wrong = [pr for pr in all_orders(oracle) if pr != oracle.PRIORITY]
    assert len(wrong) == 119
    for pr in wrong:
        assert any(rle(oracle, p, t, pen, pr)[1] != want for p, t, pen, want, _, _ in cases), pr


def test_reviewed_mutations_change_a_cigar(oracle):
    # (M, D_ext, D_open, I_ext, I_open): I before D; M last; open before extend
    for pr in [(9, 2, 1, 6, 5), (3, 6, 5, 2, 1), (9, 5, 6, 1, 2), (9, 5, 6, 2, 1), (9, 6, 5, 1, 2)]:
        for cases in (TIES_462, TIES_531):
            assert any(rle(oracle, p, t, pen, pr)[1] != want for p, t, pen, want, _, _ in cases), pr


@pytest.mark.parametrize("case", HOMOPOLYMER_462, ids=lambda c: f"{c[0]}-{c[1]}")
def test_homopolymer_gap_placement(oracle, case):
    p, t, pen, want, where, leftmost = case
    s, ops, _ = oracle.wfa(p, t, *pen, cigar=True)
    assert oracle.rle(ops) == want
    assert s == oracle.gotoh(p, t, *pen) == brute(p, t, *pen)
    assert oracle.rescore(oracle.expand(leftmost), p, t, *pen) == s
    # one candidate per step: no priority order changes it
    for pr in all_orders(oracle):
        assert rle(oracle, p, t, pen, pr)[1] == want


def test_common_prefix_only_prefixes_the_cigar(oracle):
    rng = random.Random(5)
    cases = [(p, t, pen, want) for p, t, pen, want, _, _ in ALL_TIES] + \
            [(p, t, pen, want) for p, t, pen, want, _, _ in HOMOPOLYMER_462]
    for n in (1, 9, 300, 2000, 5000):
        pre = "".join(rng.choice("ACGT") for _ in range(n))
        for p, t, pen, want in cases:
            s, ops, _ = oracle.wfa(pre + p, pre + t, *pen, cigar=True)
            assert s == oracle.wfa(p, t, *pen)[0]
            assert ops == "M" * n + oracle.expand(want), (n, p, t)
