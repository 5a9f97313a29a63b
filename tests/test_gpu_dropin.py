"""The drop-in contract on the GPU: the reference's own run script and the
meaning of wf_length.

run_new.sh:3-63 generates twelve datasets with `generate_dataset N L E` and
aligns each with `wfa_gpu 4 6 2 sequences.txt W 1`.  Every line runs here
through bin/generate_dataset + bin/wfa_gpu at a reduced N (the same lengths,
rates and wf_length values) and must print the reference's lines and its
success banner (wfa_gpu.cu:134-135,174-175,183-184) with exit status 0.
"""
import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu

# (length, error %, wf_length) of run_new.sh:6-63, in its order
RUN_NEW = [(128, 1, 21), (512, 1, 21), (1024, 1, 200), (10240, 1, 400),
           (128, 2, 42), (512, 2, 42), (1024, 2, 200), (10240, 2, 800),
           (128, 5, 105), (512, 5, 105), (1024, 5, 200), (10240, 5, 2000)]


def say(msg):
    print(f"\n[gpu] {msg}", flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("length,err,wf_length", RUN_NEW)
def test_run_new_sh_line(repo, tmp_path, length, err, wf_length):
    n = 2000 if length <= 1024 else 100
    gen = os.path.join(repo, "bin", "generate_dataset")
    r = subprocess.run([gen, str(n), str(length), str(err)], cwd=tmp_path, capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    say(f"run_new.sh line: {n} x {length} bp at {err}%, wf_length {wf_length}")
    r = subprocess.run([os.path.join(repo, "bin", "wfa_gpu"), "4", "6", "2", "sequences.txt", str(wf_length), "1"],
                       cwd=tmp_path, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    lines = r.stdout.splitlines()
    assert lines[0].startswith("GPU Time: ") and lines[1].startswith("Estimated GCUPS GPU: : ")
    assert lines[2].startswith("CPU Time: ") and lines[3].startswith("Estimated GCUPS SW: ")
    assert "|ALIGNEMENT FINISHED CORRECTLY|" in r.stdout


def test_wf_length_is_a_minimum_capacity(tsu, oracle):
    """wf_length is the least capacity the plain form uses.  The register
    kernels (4/6/2 penalties) round it up to their window (400 -> 512
    diagonals), so a pair whose range needs 403 diagonals succeeds at 400;
    the LDS kernel (other penalties) uses the capacity as given, so a pair
    whose wavefront is wider than wf_length fails there with TSU_EKOVERFLOW
    and succeeds once wf_length holds it."""
    ds = tsu.Dataset.generate(40, seed=0x5453554E, len_min=10240, err_min=0.01, edit_mode=2)
    pairs = ds.pairs()
    width = [oracle.wfa(p, t, 4, 6, 2)[2]["max_width"] for p, t in pairs]
    i = max(range(len(pairs)), key=lambda j: width[j])
    assert width[i] > 400, width[i]            # the run_new.sh line's own pairs need more than 400
    a = tsu.Aligner(4, 6, 2)
    try:
        say(f"minimum capacity: 4/6/2, need {width[i]} diagonals, wf_length 400")
        got, _ = a.align_plain(a.batch([pairs[i]]), 400)
        assert got[0] == oracle.wfa(*pairs[i], 4, 6, 2)[0]
    finally:
        a.close()
    pen = (5, 6, 2)
    s, _, c = oracle.wfa(*pairs[i], *pen)
    # a capacity of W diagonals spans [-W/2, W/2 - 1] around diagonal 0
    # (wfa_gpu.cu:217); narrower than the widest range it cannot hold it
    fits = max(-2 * c["min_lo"], 2 * c["max_hi"] + 2)
    a = tsu.Aligner(*pen)
    try:
        b = a.batch([pairs[i]])
        say(f"minimum capacity: 5/6/2, width {c['max_width']}, wf_length {c['max_width'] - 1} and {fits}")
        got, _ = a.align_plain(b, c["max_width"] - 1, raise_on_error=False)
        assert got[0] == tsu.TSU_EKOVERFLOW
        got, _ = a.align_plain(b, fits)
        assert got[0] == s
    finally:
        a.close()
