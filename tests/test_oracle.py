"""Pins the CPU oracle (oracle/wfa_oracle.c) before it is trusted.

Checks, each chosen so that a plausible mistake fails one of them (for the
gap-affine 2-piece oracle too: an explicit-run-length enumeration, closed
forms, the reduction to one piece, WFA == 2-piece Gotoh):
  - the WFA2-lib README example (published score 24, CIGAR 3M1X4M1D7M1I9M1X6M);
  - Gotoh == exhaustive enumeration of every alignment on tiny inputs;
  - closed forms (identity, one substitution, one gap of length L, empty);
  - WFA score == Gotoh on every fixture and on all 10,000 config-1 pairs;
  - CIGARs re-score to the score, consume both sequences, M/X are true
    matches/mismatches; swap symmetry; the all-mismatch-plus-gap bound;
  - the committed golden fixtures.
"""
import functools
import itertools
import json
import os
import random

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
PENALTIES = [(4, 6, 2), (1, 1, 1), (2, 3, 1), (5, 2, 3), (3, 0, 1), (1, 0, 1), (6, 4, 2)]


def brute(p, t, x, o, e):
    """Minimum cost over every alignment path (exhaustive recursion)."""
    @functools.lru_cache(None)
    def f(i, j, last):  # last: 0 none/M/X, 1 I, 2 D
        if i == len(p) and j == len(t):
            return 0
        best = 1 << 60
        if i < len(p) and j < len(t):
            best = min(best, (0 if p[i] == t[j] else x) + f(i + 1, j + 1, 0))
        if j < len(t):
            best = min(best, (e if last == 1 else o + e) + f(i, j + 1, 1))
        if i < len(p):
            best = min(best, (e if last == 2 else o + e) + f(i + 1, j, 2))
        return best
    return f(0, 0, 0)


def rand_seq(rng, n):
    return "".join(rng.choice("ACGT") for _ in range(n))


def test_wfa2lib_readme_example(oracle):
    p = "TCTTTACTCGCGCGTTGGAGAAATACAATAGT"
    t = "TCTATACTGCGCGTTTGGAGAAATAAAATAGT"
    s, ops, _ = oracle.wfa(p, t, 4, 6, 2, cigar=True)
    assert s == 24
    assert oracle.rle(ops) == "3M1X4M1D7M1I9M1X6M"
    assert oracle.gotoh(p, t, 4, 6, 2) == 24


@pytest.mark.parametrize("pen", PENALTIES)
def test_gotoh_and_wfa_equal_brute_force(oracle, pen):
    rng = random.Random(hash(pen) & 0xFFFF)
    for _ in range(400):
        p = rand_seq(rng, rng.randint(0, 7))
        t = rand_seq(rng, rng.randint(0, 7))
        b = brute(p, t, *pen)
        assert oracle.gotoh(p, t, *pen) == b, (p, t)
        s, ops, _ = oracle.wfa(p, t, *pen, cigar=True)
        assert s == b, (p, t, pen)
        assert oracle.rescore(ops, p, t, *pen) == b


def test_closed_forms(oracle):
    rng = random.Random(3)
    for x, o, e in PENALTIES:
        for L in (1, 5, 37, 100):
            s = rand_seq(rng, 60)
            assert oracle.wfa(s, s, x, o, e)[0] == 0
            # one substitution costs min(x, 2(o+e)) (a sub can also be a D+I pair)
            i = rng.randrange(len(s))
            sub = s[:i] + "ACGT"[("ACGT".index(s[i]) + 1) % 4] + s[i + 1:]
            assert oracle.wfa(s, sub, x, o, e)[0] == min(x, 2 * (o + e))
            # empty against L bases: one gap
            t = rand_seq(rng, L)
            assert oracle.wfa("", t, x, o, e)[0] == o + L * e
            assert oracle.wfa(t, "", x, o, e)[0] == o + L * e
        assert oracle.wfa("", "", x, o, e)[0] == 0


def test_single_indel_closed_form(oracle):
    # a homopolymer-free insertion of L bases into a sequence costs o + L*e
    base = "ACGTTGCAAGCTTCGA" * 4
    for L in (1, 2, 3, 8):
        ins = base[:20] + "C" * L + base[20:]
        s, ops, _ = oracle.wfa(base, ins, 4, 6, 2, cigar=True)
        assert s == 6 + 2 * L
        assert oracle.rescore(ops, base, ins, 4, 6, 2) == s


def test_gcd_divisibility_and_symmetry(oracle):
    rng = random.Random(11)
    for _ in range(200):
        p = rand_seq(rng, rng.randint(0, 60))
        t = list(p)
        for _ in range(rng.randint(0, 10)):
            pos = rng.randrange(len(t) + 1)
            op = rng.randrange(3)
            if op == 0 and pos < len(t):
                t[pos] = rng.choice("ACGT")
            elif op == 1:
                t.insert(pos, rng.choice("ACGT"))
            elif pos < len(t):
                del t[pos]
        t = "".join(t)
        s = oracle.wfa(p, t, 4, 6, 2)[0]
        assert s % 2 == 0
        assert oracle.wfa(t, p, 4, 6, 2)[0] == s  # I <-> D symmetric
        bound = 4 * min(len(p), len(t)) + (6 + 2 * abs(len(p) - len(t)) if len(p) != len(t) else 0)
        assert s <= bound


def test_config1_all_pairs_wfa_equals_gotoh(oracle, tsu):
    ds = tsu.Dataset.generate(10000, seed=0x5453554E + 1, len_min=100, err_min=0.02)
    pairs = ds.pairs()
    scores, _, cnt = oracle.wfa_batch(pairs, 4, 6, 2)
    for i in range(0, len(pairs)):
        p, t = pairs[i]
        assert scores[i] == oracle.gotoh(p, t, 4, 6, 2), i
    assert cnt["cells"] > 0 and cnt["extend_words"] > 0


def test_golden_fixtures(oracle):
    with open(os.path.join(HERE, "golden", "pairs_scores_cigars.json")) as f:
        cases = json.load(f)["cases"]
    assert len(cases) > 300
    for c in cases:
        pen = tuple(c["penalties"])
        assert oracle.gotoh(c["pattern"], c["text"], *pen) == c["score"], c["name"]
        s, ops, _ = oracle.wfa(c["pattern"], c["text"], *pen, cigar=True)
        assert s == c["score"], c["name"]
        assert oracle.rle(ops) == c["cigar"], c["name"]


def test_batch_matches_single_and_cigars_valid(oracle, tsu):
    ds = tsu.Dataset.generate(300, seed=9, len_min=80, len_max=900, err_min=0.0, err_max=0.2)
    pairs = ds.pairs()
    scores, ops, _ = oracle.wfa_batch(pairs, 4, 6, 2, threads=4, cigar=True)
    for i, (p, t) in enumerate(pairs):
        s1, o1, _ = oracle.wfa(p, t, 4, 6, 2, cigar=True)
        assert scores[i] == s1
        assert ops[i] == o1
        assert oracle.rescore(ops[i], p, t, 4, 6, 2) == s1


def test_rescore_catches_a_planted_mistake(oracle):
    # a plausible backtrace bug (one op dropped / swapped) must be detected
    p = "TCTTTACTCGCGCGTTGGAGAAATACAATAGT"
    t = "TCTATACTGCGCGTTTGGAGAAATAAAATAGT"
    _, ops, _ = oracle.wfa(p, t, 4, 6, 2, cigar=True)
    bad = ops.replace("D", "", 1)
    with pytest.raises(Exception):
        oracle.rescore(bad, p, t, 4, 6, 2)
    swapped = ops.replace("I", "Q", 1).replace("D", "I", 1).replace("Q", "D", 1)
    with pytest.raises(Exception):
        oracle.rescore(swapped, p, t, 4, 6, 2)


# ---------------------------------------------------------------------------
# gap-affine 2-piece (oracle/wfa2p_oracle.c)
# ---------------------------------------------------------------------------
PENALTIES_2P = [(4, 6, 2, 24, 1), (4, 6, 2, 12, 1), (2, 3, 2, 8, 1), (5, 1, 3, 10, 1), (3, 2, 2, 2, 2),
                (1, 0, 1, 4, 1), (6, 4, 3, 9, 1)]


def brute2p(p, t, x, o1, e1, o2, e2):
    """Minimum over alignments written as runs: each gap run of length L costs
    min(o1 + L e1, o2 + L e2) (explicit run lengths, no gap state machine)."""
    g = lambda L: min(o1 + L * e1, o2 + L * e2)

    @functools.lru_cache(None)
    def f(i, j, after_gap):  # after_gap: 0 after a match/mismatch or at start, 1 after an I run, 2 after a D run
        if i == len(p) and j == len(t):
            return 0
        best = 1 << 60
        if i < len(p) and j < len(t):
            best = min(best, (0 if p[i] == t[j] else x) + f(i + 1, j + 1, 0))
        if after_gap != 1:
            for L in range(1, len(t) - j + 1):
                best = min(best, g(L) + f(i, j + L, 1))
        if after_gap != 2:
            for L in range(1, len(p) - i + 1):
                best = min(best, g(L) + f(i + L, j, 2))
        return best
    return f(0, 0, 0)


@pytest.mark.parametrize("pen", PENALTIES_2P)
def test_2p_gotoh_and_wfa_equal_brute_force(oracle, pen):
    rng = random.Random(sum(pen) * 7919)
    for _ in range(250):
        p = rand_seq(rng, rng.randint(0, 7))
        t = rand_seq(rng, rng.randint(0, 7))
        b = brute2p(p, t, *pen)
        assert oracle.gotoh2p(p, t, *pen) == b, (p, t, pen)
        assert oracle.wfa2p(p, t, *pen)[0] == b, (p, t, pen)


def test_2p_closed_forms_and_reductions(oracle):
    rng = random.Random(11)
    for x, o1, e1, o2, e2 in PENALTIES_2P:
        pen = (x, o1, e1, o2, e2)
        s = rand_seq(rng, 80)
        assert oracle.wfa2p(s, s, *pen)[0] == 0
        assert oracle.wfa2p("", "", *pen)[0] == 0
        for L in (1, 3, 9, 40):
            gap = min(o1 + L * e1, o2 + L * e2)
            assert oracle.wfa2p("", s[:L], *pen)[0] == gap
            # one extra run of length L inside an otherwise identical pair: any
            # alignment needs gap runs summing to L, which cost at least g(L)
            assert oracle.wfa2p(s[:20] + "A" * L + s[20:], s, *pen)[0] == gap
            assert oracle.wfa2p(s, s[:20] + "C" * L + s[20:], *pen)[0] == gap
            # a single gap piece (both pieces equal) reduces to gap-affine
            a = rand_seq(rng, 50)
            b = a[:10] + rand_seq(rng, L) + a[10 + L // 2:]
            assert oracle.wfa2p(a, b, x, o1, e1, o1, e1)[0] == oracle.gotoh(a, b, x, o1, e1)
        # the second piece can only lower the score
        a, b = rand_seq(rng, 60), rand_seq(rng, 70)
        assert oracle.wfa2p(a, b, *pen)[0] <= oracle.gotoh(a, b, x, o1, e1)
        assert oracle.wfa2p(a, b, *pen)[0] == oracle.wfa2p(b, a, *pen)[0]


def test_2p_wfa_equals_gotoh_on_longer_pairs(oracle, tsu):
    ds = tsu.Dataset.generate(120, seed=0x2F, len_min=200, len_max=1500, err_min=0.02, err_max=0.2)
    pairs = ds.pairs()
    for pen in ((4, 6, 2, 24, 1), (5, 1, 3, 10, 1)):
        got, _ = oracle.wfa2p_batch(pairs, *pen)
        want = [oracle.gotoh2p(p, t, *pen) for p, t in pairs[:40]]
        assert list(got[:40]) == want, pen


@pytest.mark.parametrize("pen", [(2, 3, 1), (4, 6, 2), (1, 1, 1), (3, 5, 1)])
def test_nulled_out_of_bounds_equals_trimmed(oracle, pen):
    """The neighbour-synchronised GPU kernel NULLs every out-of-bounds I, D
    and M cell where it is computed instead of trimming the wavefront ends
    (it has no whole-window view per score).  Its score and its CIGAR (walked
    over the NULLed cells with the same priority) equal the trimmed WFA's on
    every pair of up to 3 bases and on ragged random pairs (unequal lengths
    push cells out of bounds at both ends); the score equals Gotoh's."""
    alphabet = "ACGT"
    seqs = [""]
    for n in range(1, 4):
        seqs += ["".join(c) for c in itertools.product(alphabet, repeat=n)]
    for p in seqs:
        for t in seqs:
            s, ops, _ = oracle.wfa(p, t, *pen, cigar=True)
            assert oracle.wfa_nulled(p, t, *pen, cigar=True) == (s, ops), (p, t)
    rng = random.Random(0x4E55 + sum(pen))
    for _ in range(1500):
        p = rand_seq(rng, rng.randint(1, 150))
        mode = rng.randrange(3)
        if mode == 0:
            t = rand_seq(rng, rng.randint(1, 150))
        else:
            t = list(p)
            for _ in range(rng.randint(0, len(p) // 3)):
                i = rng.randrange(len(t) + 1)
                r = rng.randrange(3)
                if r == 0 and i < len(t):
                    t[i] = rng.choice([c for c in alphabet if c != t[i]])
                elif r == 1:
                    t.insert(i, rng.choice(alphabet))
                elif i < len(t):
                    del t[i]
            t = "".join(t) + (rand_seq(rng, rng.randint(0, 40)) if mode == 2 else "")
        s, ops, _ = oracle.wfa(p, t, *pen, cigar=True)
        assert oracle.wfa_nulled(p, t, *pen, cigar=True) == (s, ops), (p, t)
    for _ in range(100):
        p, t = rand_seq(rng, rng.randint(1, 60)), rand_seq(rng, rng.randint(1, 60))
        assert oracle.wfa_nulled(p, t, *pen) == oracle.gotoh(p, t, *pen)
