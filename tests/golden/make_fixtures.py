"""Writes the committed golden fixtures of this directory.

    python tests/golden/make_fixtures.py

Pairs are produced by the library's seeded generator (pure host code) and by
hand; expected scores come from the oracle's Gotoh DP (textbook, independent
of any WFA code) and are cross-checked against the oracle's WFA; expected
CIGARs come from the oracle's WFA backtrace and are checked to re-score to
the Gotoh score.  The WFA2-lib README example (published by smarco/WFA2-lib)
is included with its published score and CIGAR.
"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, REPO)

from oracle import oracle as orc  # noqa: E402
import tsunami_pact_amd as tsu  # noqa: E402

PEN = (4, 6, 2)


def case(p, t, pen=PEN, name=""):
    g = orc.gotoh(p, t, *pen)
    s, ops, _ = orc.wfa(p, t, *pen, cigar=True)
    assert s == g, (p, t, pen, s, g)
    assert orc.rescore(ops, p, t, *pen) == g
    return {"name": name, "pattern": p, "text": t, "penalties": list(pen), "score": g, "cigar": orc.rle(ops)}


def main():
    cases = []
    # WFA2-lib README example: published score -24 and CIGAR
    ex = case("TCTTTACTCGCGCGTTGGAGAAATACAATAGT", "TCTATACTGCGCGTTTGGAGAAATAAAATAGT", name="wfa2lib_readme")
    assert ex["score"] == 24 and ex["cigar"] == "3M1X4M1D7M1I9M1X6M"
    cases.append(ex)
    hand = [
        ("", ""), ("", "ACGT"), ("ACGT", ""), ("A", "A"), ("A", "C"), ("ACGT", "ACGT"),
        ("ACGTACGT", "ACGAACGT"), ("AAAA", "AAAAAA"), ("AAAAAA", "AAAA"), ("ACGT" * 8, "TGCA" * 8),
        ("A" * 40, "C" * 40), ("ACGTTGCA" * 5, "ACGTTGCA" * 5 + "GG"), ("GATTACA", "GCATGCT"),
    ]
    for i, (p, t) in enumerate(hand):
        for pen in (PEN, (1, 1, 1), (2, 3, 1), (5, 2, 3), (3, 0, 1)):
            cases.append(case(p, t, pen, name=f"hand{i}"))
    # config-1 shape: 100 bp, 2% mixed edits
    ds = tsu.Dataset.generate(200, seed=0x5453554E + 1, len_min=100, err_min=0.02)
    for i, (p, t) in enumerate(ds.pairs()):
        cases.append(case(p, t, name=f"cfg1_{i}"))
    # mixed lengths and rates
    ds = tsu.Dataset.generate(60, seed=77, len_min=50, len_max=600, err_min=0.01, err_max=0.2)
    for i, (p, t) in enumerate(ds.pairs()):
        cases.append(case(p, t, name=f"mixed_{i}"))
    with open(os.path.join(HERE, "pairs_scores_cigars.json"), "w") as f:
        json.dump({"generator": "tests/golden/make_fixtures.py", "cases": cases}, f, indent=0)
    # generator golden: first pairs of fixed seeds (pins the generator's output)
    gen = {}
    for name, kw in {
        "cfg1": dict(n=4, seed=0x5453554E + 1, len_min=100, err_min=0.02),
        "cfg2": dict(n=2, seed=0x5453554E + 2, len_min=1000, err_min=0.05),
        "cfg4": dict(n=4, seed=0x5453554E + 4, len_min=100, len_max=20000, err_min=0.01, err_max=0.15),
        "sub": dict(n=3, seed=5, len_min=64, err_min=0.05, edit_mode=1),
    }.items():
        d = tsu.Dataset.generate(**kw)
        gen[name] = {"args": kw, "pairs": d.pairs()}
    with open(os.path.join(HERE, "generator.json"), "w") as f:
        json.dump(gen, f, indent=0)
    print(f"{len(cases)} cases written")


if __name__ == "__main__":
    main()
