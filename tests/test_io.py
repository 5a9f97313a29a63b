"""Host-side inputs: generator, sequences.txt reader/writer, 2-bit packer.

The reference format is read by wfa_gpu.cu:71-81 and written by
generate_dataset.cpp:128-158; the generator follows the seeded, indel-mixed
specification of the benchmark configurations.
"""
import json
import os
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


def levenshtein_ops_bound(p, t):
    return abs(len(p) - len(t))


def test_generator_is_deterministic_and_thread_independent(tsu, tmp_path):
    a = tsu.Dataset.generate(3000, seed=123, len_min=100, len_max=2000, err_min=0.01, err_max=0.15)
    b = tsu.Dataset.generate(3000, seed=123, len_min=100, len_max=2000, err_min=0.01, err_max=0.15)
    assert a.pairs() == b.pairs()
    env = dict(os.environ, OMP_NUM_THREADS="1")
    code = ("import sys; sys.path.insert(0, %r); import tsunami_pact_amd as t; import json;"
            "d = t.Dataset.generate(3000, seed=123, len_min=100, len_max=2000, err_min=0.01, err_max=0.15);"
            "print(json.dumps(d.pairs()[::97]))") % os.path.dirname(HERE)
    out = subprocess.run(["python", "-c", code], env=env, capture_output=True, text=True, check=True).stdout
    assert [tuple(x) for x in json.loads(out)] == a.pairs()[::97]
    c = tsu.Dataset.generate(3000, seed=124, len_min=100, len_max=2000, err_min=0.01, err_max=0.15)
    assert c.pairs() != a.pairs()


def test_generator_golden(tsu):
    with open(os.path.join(HERE, "golden", "generator.json")) as f:
        gold = json.load(f)
    for name, g in gold.items():
        d = tsu.Dataset.generate(**g["args"])
        assert [list(x) for x in d.pairs()] == g["pairs"], name


def test_generator_edit_model(tsu):
    # config-1 shape: 100 bp, exactly round(0.02*100) = 2 edits per pair
    d = tsu.Dataset.generate(4000, seed=1, len_min=100, err_min=0.02)
    pl, tl = d.lengths()
    assert (pl == 100).all()
    # an insertion lengthens, a deletion shortens: |t|-|p| in [-2, 2]
    diff = tl.astype(int) - pl
    assert diff.min() >= -2 and diff.max() <= 2
    # the three edit kinds are used in roughly equal proportion
    counts = {k: int((diff == k).sum()) for k in range(-2, 3)}
    assert counts[0] > 0 and counts[1] > 0 and counts[-1] > 0
    # substitution-only mode keeps lengths and changes exactly round(rate*L) bases
    s = tsu.Dataset.generate(500, seed=2, len_min=200, err_min=0.05, edit_mode=1)
    for p, t in s.pairs():
        assert len(p) == len(t) == 200
        assert sum(a != b for a, b in zip(p, t)) == 10
    # the reference's count, len * percent / 100 rounded down: 10240 at 2% is
    # 204 substitutions where rounding would give 205 (generate_dataset.cpp:23)
    r = tsu.Dataset.generate(20, seed=3, len_min=10240, err_min=0.02, edit_mode=2)
    for p, t in r.pairs():
        assert len(p) == len(t) == 10240
        assert sum(a != b for a, b in zip(p, t)) == 204
    assert sum(a != b for a, b in zip(*tsu.Dataset.generate(1, seed=3, len_min=10240, err_min=0.02,
                                                             edit_mode=1).pair(0))) == 205


def test_generator_length_and_rate_ranges(tsu):
    d = tsu.Dataset.generate(5000, seed=4, len_min=100, len_max=20000, err_min=0.01, err_max=0.15)
    pl, _ = d.lengths()
    assert pl.min() >= 100 and pl.max() <= 20000
    # log-uniform: about half the pairs are below the geometric mean sqrt(100*20000)=1414
    frac = (pl < 1414).mean()
    assert 0.45 < frac < 0.55


def test_write_read_roundtrip(tsu, tmp_path):
    d = tsu.Dataset.generate(200, seed=5, len_min=50, len_max=500, err_min=0.0, err_max=0.2)
    path = str(tmp_path / "sequences.txt")
    d.write(path, 7)
    r = tsu.Dataset.read(path)
    assert r.pairs() == d.pairs()
    num, L, err = r.header()
    assert num == 200 and err == 7
    assert L == max(max(len(p), len(t)) for p, t in d.pairs())


def test_reader_fixed_slot_truncation(tsu, tmp_path):
    # tokens longer than the header length keep their first len characters,
    # as the reference's fixed slots do (wfa_gpu.cu:78-81)
    path = tmp_path / "legacy.txt"
    path.write_text("2\n8\n1\nACGTACGT1\nACGTACGA\xe1\nACGTACGTGG\nTTTTTTTT\n", encoding="latin-1")
    r = tsu.Dataset.read(str(path))
    assert r.pairs() == [("ACGTACGT", "ACGTACGA"), ("ACGTACGT", "TTTTTTTT")]


@pytest.mark.parametrize("content,msg", [
    ("", "truncated header"),
    ("2\n8\n", "truncated header"),
    ("x\n8\n1\n", "not an integer"),
    ("2\n4\n1\nACGT\nACGT\nACGT\n", "file ends at pair 1"),
    ("1\n4\n1\nACNT\nACGT\n", "non-ACGT"),
])
def test_reader_errors(tsu, tmp_path, content, msg):
    path = tmp_path / "bad.txt"
    path.write_text(content)
    with pytest.raises(tsu.TsuError) as ei:
        tsu.Dataset.read(str(path))
    assert ei.value.status == tsu.TSU_EBADINPUT
    assert msg in str(ei.value)


def test_reader_missing_file(tsu, tmp_path):
    with pytest.raises(tsu.TsuError) as ei:
        tsu.Dataset.read(str(tmp_path / "nope.txt"))
    assert ei.value.status == tsu.TSU_EIO


def test_reads_reference_generator_output(tsu, tmp_path, oracle):
    """The reference's own generate_dataset (compiled from its source into
    oracle/_ref) writes files the product reads; L = 40 exercises the
    missing-NUL garbage after each token (generate_dataset.cpp:115-116)."""
    if not os.path.exists(oracle.REF_GENERATOR):
        pytest.skip("oracle/_ref/generate_dataset not built (reference tree absent)")
    for L, E in ((40, 5), (100, 2), (128, 1)):
        subprocess.run([oracle.REF_GENERATOR, "50", str(L), str(E)], cwd=tmp_path, check=True,
                       capture_output=True)
        d = tsu.Dataset.read(str(tmp_path / "sequences.txt"))
        assert len(d) == 50
        assert d.header()[:2] == (50, L)
        for p, t in d.pairs():
            assert len(p) == len(t) == L
            # substitution-only: Hamming distance equals the reference's error count
            assert sum(a != b for a, b in zip(p, t)) <= max(1, L * E // 100)
        sc, _, _ = oracle.wfa_batch(d.pairs(), 4, 6, 2)
        assert all(sc[i] == oracle.gotoh(p, t, 4, 6, 2) for i, (p, t) in enumerate(d.pairs()))


def _pack_py(seq):
    code = {"A": 0, "C": 1, "G": 2, "T": 3}
    n = (len(seq) + 15) // 16
    out = np.zeros(n, np.uint32)
    for i, c in enumerate(seq):
        out[i // 16] |= np.uint32(code[c] << (2 * (i % 16)))
    return out


def test_pack_layout(tsu):
    rng = np.random.default_rng(0)
    for L in (0, 1, 15, 16, 17, 31, 32, 33, 100, 1000):
        s = "".join(rng.choice(list("ACGT"), L))
        assert (tsu.pack(s) == _pack_py(s)).all()
    with pytest.raises(tsu.TsuError):
        tsu.pack("ACGN")
    with pytest.raises(tsu.TsuError):
        tsu.pack("acgt")
