"""The drop-in CLIs keep the reference's argument contract.

wfa_gpu: exactly six arguments, else the reference's message and exit 0
(wfa_gpu.cu:45-48); a missing file prints "ERROR: Cannot open file." and
exits 1 (wfa_gpu.cu:66-69).  generate_dataset: fewer than three arguments
prints the reference's usage and exits 1 (generate_dataset.cpp:14-18).
"""
import os
import subprocess

import pytest


def run(args, cwd=None, env=None):
    return subprocess.run(args, capture_output=True, text=True, cwd=cwd, env=env, timeout=120)


@pytest.fixture(scope="module")
def bins(repo):
    w = os.path.join(repo, "bin", "wfa_gpu")
    g = os.path.join(repo, "bin", "generate_dataset")
    assert os.path.exists(w) and os.path.exists(g), "build first: make -C tsunami_pact_amd"
    return w, g


def test_wfa_gpu_wrong_argc(bins):
    for args in ([], ["4", "6", "2"], ["4", "6", "2", "f", "100", "1", "extra"]):
        r = run([bins[0]] + args)
        assert r.returncode == 0
        assert r.stdout == ("ERROR! Please specify in order: mismatch, gap_opening, gap_extension, "
                            "file name, wf_length, check\n")


def test_wfa_gpu_missing_file(bins, tmp_path):
    r = run([bins[0], "4", "6", "2", str(tmp_path / "none.txt"), "100", "1"])
    assert r.returncode == 1
    assert r.stdout == "ERROR: Cannot open file.\n"


def test_wfa_gpu_bad_file_fails_loudly(bins, tmp_path):
    p = tmp_path / "bad.txt"
    p.write_text("1\n4\n1\nACNT\nACGT\n")
    r = run([bins[0], "4", "6", "2", str(p), "100", "0"])
    assert r.returncode != 0
    assert "non-ACGT" in r.stdout


def test_generate_dataset_usage(bins):
    r = run([bins[1]])
    assert r.returncode == 1
    assert r.stdout.strip() == "Required arguments: number of couples, sequences length, error rate"


def test_generate_dataset_seeded(bins, tmp_path, tsu):
    """With the reference's three arguments (plus a seed) the file holds the
    reference generator's model: equal lengths, substitutions only,
    len * err / 100 rounded down of them per pair (generate_dataset.cpp:
    20-23,45-66); "mixed" switches to sub/ins/del edits."""
    a = tmp_path / "a"; b = tmp_path / "b"
    a.mkdir(); b.mkdir()
    for d in (a, b):
        r = run([bins[1], "100", "128", "2", "42"], cwd=d)
        assert r.returncode == 0
    ta = (a / "sequences.txt").read_text()
    assert ta == (b / "sequences.txt").read_text()
    lines = ta.split("\n")
    assert lines[0] == "100" and lines[1] == "128" and lines[2] == "2"
    ds = tsu.Dataset.read(str(a / "sequences.txt"))
    assert len(ds) == 100
    assert ds.pairs() == tsu.Dataset.generate(100, seed=42, len_min=128, err_min=0.02, edit_mode=2).pairs()
    for p, t in ds.pairs():
        assert len(p) == len(t) == 128
        assert sum(x != y for x, y in zip(p, t)) == 2   # 128 * 2 // 100
    # the default seed and mode, exactly the reference's argument list
    r = run([bins[1], "10", "64", "5"], cwd=b)
    assert r.returncode == 0
    for p, t in tsu.Dataset.read(str(b / "sequences.txt")).pairs():
        assert len(p) == len(t) == 64
        assert sum(x != y for x, y in zip(p, t)) == 3   # 64 * 5 // 100
    r = run([bins[1], "10", "64", "5", "7", "sub"], cwd=a)
    assert r.returncode == 0
    assert tsu.Dataset.read(str(a / "sequences.txt")).pairs() == \
        tsu.Dataset.generate(10, seed=7, len_min=64, err_min=0.05, edit_mode=2).pairs()
    r = run([bins[1], "100", "128", "2", "42", "mixed"], cwd=a)
    assert r.returncode == 0
    assert tsu.Dataset.read(str(a / "sequences.txt")).pairs() == \
        tsu.Dataset.generate(100, seed=42, len_min=128, err_min=0.02).pairs()
    r = run([bins[1], "10", "64", "5", "7", "indels"], cwd=a)
    assert r.returncode == 1
