"""GPU parity: libtsunami_hip.so (HIP, gfx950) vs the CPU oracle, bit-exact.

Every test calls the C-ABI through tsunami_pact_amd (ctypes) and compares
integer scores (and CIGARs) with oracle/ on the same seeded inputs.  Sizes
are those the oracle finishes in seconds; the full-size configuration is
checked through size-independent properties plus a seeded sample.
"""
import json
import os
import subprocess
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
SEED = 0x5453554E


def say(msg):
    print(f"\n[gpu] {msg}", flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def fixtures():
    with open(os.path.join(HERE, "golden", "pairs_scores_cigars.json")) as f:
        return json.load(f)["cases"]


@pytest.fixture(scope="module")
def aligners(tsu):
    cache = {}

    def get(pen):
        if pen not in cache:
            cache[pen] = tsu.Aligner(*pen)
        return cache[pen]
    yield get
    for a in cache.values():
        a.close()


def test_device_is_gfx950(aligner):
    info = aligner.device_info()
    say(f"device {info}")
    assert "gfx950" in info["name"]
    assert info["cu_count"] > 0


def test_plain_config1_all_pairs(aligner, tsu, oracle):
    """Config 1: 10,000 pairs of 100 bp, 2% mixed edits, W = 100."""
    ds = tsu.Dataset.generate(10000, seed=SEED + 1, len_min=100, err_min=0.02)
    say("plain config1 10000x100 W=100")
    b = aligner.batch(ds)
    got, t = aligner.align_plain(b, 100)
    want, _, _ = oracle.wfa_batch(ds.pairs(), 4, 6, 2)
    assert (got == want).all(), np.nonzero(got != want)[0][:10]
    assert t["kernel_ms"] > 0


def test_plain_config2_sample(aligner, tsu, oracle):
    """Config-2 shape: 1000 bp, 5% mixed edits (3000 seeded pairs)."""
    ds = tsu.Dataset.generate(3000, seed=SEED + 2, len_min=1000, err_min=0.05)
    want, _, cnt = oracle.wfa_batch(ds.pairs(), 4, 6, 2)
    W = 2 * ((int(want.max()) - 6) // 2) + 3
    say(f"plain config2 sample 3000x1000 W={W}")
    b = aligner.batch(ds)
    got, _ = aligner.align_plain(b, W)
    assert (got == want).all()


def test_plain_long_reads(aligner, tsu, oracle):
    """Config-3 shape: 10,000 bp at 10% (12 seeded pairs), ring of ~7k diagonals."""
    ds = tsu.Dataset.generate(12, seed=SEED + 3, len_min=10000, err_min=0.10)
    want, _, _ = oracle.wfa_batch(ds.pairs(), 4, 6, 2)
    W = 2 * ((int(want.max()) - 6) // 2) + 3
    say(f"plain long 12x10000 W={W}")
    got, _ = aligner.align_plain(aligner.batch(ds), W)
    assert (got == want).all()


def test_capacity_variants_agree(aligner, tsu, oracle):
    """Different diagonal capacities select 1..16 waves per pair; same scores."""
    ds = tsu.Dataset.generate(400, seed=SEED + 5, len_min=600, err_min=0.08)
    want, _, _ = oracle.wfa_batch(ds.pairs(), 4, 6, 2)
    b = aligner.batch(ds)
    for W in (400, 700, 1300, 2500, 5000):
        say(f"capacity W={W}")
        got, _ = aligner.align_plain(b, W)
        assert (got == want).all(), W


def test_golden_fixture_scores(aligners, tsu, fixtures):
    by_pen = {}
    for c in fixtures:
        by_pen.setdefault(tuple(c["penalties"]), []).append(c)
    for pen, cases in by_pen.items():
        say(f"fixtures penalties={pen} n={len(cases)}")
        a = aligners(pen)
        b = a.batch([(c["pattern"], c["text"]) for c in cases])
        W = max(len(c["pattern"]) + len(c["text"]) + 3 for c in cases)
        got, _ = a.align_plain(b, W)
        assert [int(v) for v in got] == [c["score"] for c in cases]


def test_penalty_sets_random(aligners, tsu, oracle):
    ds = tsu.Dataset.generate(600, seed=SEED + 6, len_min=30, len_max=700, err_min=0.0, err_max=0.25)
    pairs = ds.pairs()
    for pen in [(1, 1, 1), (2, 3, 1), (5, 2, 3), (3, 0, 1), (6, 4, 2), (7, 11, 3)]:
        say(f"random penalties={pen}")
        want, _, _ = oracle.wfa_batch(pairs, *pen)
        a = aligners(pen)
        got, _ = a.align_plain(a.batch(ds), 1500)
        assert (got == want).all(), pen


def test_overflow_is_reported_not_corrupting(aligner, tsu, oracle):
    good = tsu.Dataset.generate(50, seed=1, len_min=200, err_min=0.01).pairs()
    rng = np.random.default_rng(2)
    bad = ("".join(rng.choice(list("ACGT"), 300)), "".join(rng.choice(list("ACGT"), 300)))
    pairs = good[:25] + [bad] + good[25:]
    say("overflow W=24")
    b = aligner.batch(pairs)
    got, _ = aligner.align_plain(b, 24, raise_on_error=False)
    assert got[25] == tsu.TSU_EKOVERFLOW
    want, _, _ = oracle.wfa_batch(pairs, 4, 6, 2)
    ok = [i for i in range(len(pairs)) if i != 25]
    assert (got[ok] == want[ok]).all()
    with pytest.raises(tsu.TsuError) as ei:
        aligner.align_plain(b, 24)
    assert ei.value.status == tsu.TSU_EKOVERFLOW


def test_int32_offsets_long_pair(aligner, tsu, oracle):
    """Sequences longer than 32,000 bases switch the ring to int32 offsets."""
    ds = tsu.Dataset.generate(2, seed=SEED + 7, len_min=40000, err_min=0.004)
    want, _, _ = oracle.wfa_batch(ds.pairs(), 4, 6, 2)
    W = 2 * ((int(want.max()) - 6) // 2) + 3
    say(f"int32 2x40000 W={W}")
    got, _ = aligner.align_plain(aligner.batch(ds), W)
    assert (got == want).all()


def test_single_wave_kernels_short_and_long_sequences(aligner, tsu, oracle):
    """Windows of at most 512 diagonals run on one wave per pair: with 8
    chunks and sequences up to 4,096 bases the older M scores sit in an int16
    LDS ring, otherwise the whole M ring is in registers.  Both give the
    oracle's scores."""
    short = tsu.Dataset.generate(200, seed=SEED + 17, len_min=1000, len_max=4000, err_min=0.005, err_max=0.02)
    long_ = tsu.Dataset.generate(6, seed=SEED + 18, len_min=17000, len_max=30000, err_min=0.001, err_max=0.002)
    for name, ds in (("<= 4096", short), ("> 16000", long_)):
        want, _, _ = oracle.wfa_batch(ds.pairs(), 4, 6, 2)
        say(f"single wave {name} W=512")
        got, _ = aligner.align_plain(aligner.batch(ds), 512)
        assert (got == want).all(), name


def test_sequence_limit_2p_is_lds(tsu):
    """The 2-piece kernel stages both sequences in LDS: a 2-piece context
    refuses longer ones when the batch is made, with TSU_EINVAL and a message
    (1-piece contexts read them from HBM: test_gpu_long_sequences.py)."""
    huge = tsu.Dataset.generate(1, seed=5, len_min=400_000, err_min=0.0)
    a = tsu.Aligner(4, 6, 2, gap_opening2=24, gap_extension2=1)
    try:
        say("2-piece: refuse 400 kbp")
        with pytest.raises(tsu.TsuError) as ei:
            a.batch(huge)
        assert ei.value.status == tsu.TSU_EINVAL and "LDS" in str(ei.value)
    finally:
        a.close()


def test_fixed_layout_entry_point(aligner, tsu, oracle):
    """tsu_align_fixed takes the reference's fixed-slot char arrays."""
    ds = tsu.Dataset.generate(500, seed=SEED + 8, len_min=128, err_min=0.02, edit_mode=1)
    pairs = ds.pairs()
    say("fixed layout 500x128")
    got, _ = aligner.align_fixed([p for p, _ in pairs], [t for _, t in pairs], 128, 42)
    want, _, _ = oracle.wfa_batch(pairs, 4, 6, 2)
    assert (got == want).all()


def test_batched_matches_oracle_mixed(aligner, tsu, oracle):
    """Config-4 shape (small): lengths log-uniform 100..5000, 1-15% error."""
    ds = tsu.Dataset.generate(1500, seed=SEED + 4, len_min=100, len_max=5000, err_min=0.01, err_max=0.15)
    want, _, _ = oracle.wfa_batch(ds.pairs(), 4, 6, 2)
    say("batched mixed 1500")
    got, t = aligner.align_batched(aligner.batch(ds))
    assert (got == want).all()
    assert t["launches"] >= 1


def test_cigars_match_oracle(aligner, tsu, oracle, fixtures):
    ds = tsu.Dataset.generate(300, seed=SEED + 9, len_min=50, len_max=3000, err_min=0.0, err_max=0.15)
    pairs = ds.pairs()
    say("cigar mixed 300")
    got, cig, _ = aligner.align_cigar(aligner.batch(ds))
    want, ops, _ = oracle.wfa_batch(pairs, 4, 6, 2, cigar=True)
    assert (got == want).all()
    for i, (p, t) in enumerate(pairs):
        assert cig[i] == oracle.rle(ops[i]), i
        assert oracle.rescore(oracle.expand(cig[i]), p, t, 4, 6, 2) == got[i]
    # the run count over all pairs: one run per letter group of each CIGAR
    _, _, tc = aligner.align_cigar(aligner.batch(ds), strings=False, count_runs=True)
    assert tc["cigar_runs"] == sum(sum(1 for ch in c if ch.isalpha()) for c in cig)
    cases = [c for c in fixtures if tuple(c["penalties"]) == (4, 6, 2)]
    say(f"cigar fixtures {len(cases)}")
    got, cig, _ = aligner.align_cigar(aligner.batch([(c["pattern"], c["text"]) for c in cases]))
    assert [int(v) for v in got] == [c["score"] for c in cases]
    assert cig == [c["cigar"] for c in cases]


def test_cigar_wfa2lib_readme_example(aligner):
    say("cigar readme example")
    got, cig, _ = aligner.align_cigar(aligner.batch([("TCTTTACTCGCGCGTTGGAGAAATACAATAGT",
                                                        "TCTATACTGCGCGTTTGGAGAAATAAAATAGT")]))
    assert int(got[0]) == 24
    assert cig[0] == "3M1X4M1D7M1I9M1X6M"


def test_full_size_config2_properties(aligner, tsu, oracle):
    """1,000,000 pairs of 1000 bp at 5%: every score is non-negative, a
    multiple of gcd(4,6,2)=2 and within the all-mismatch-plus-gap bound; a
    seeded sample of 2,000 pairs equals the oracle; re-running is identical."""
    n = 1_000_000
    ds = tsu.Dataset.generate(n, seed=SEED + 2, len_min=1000, err_min=0.05)
    say("full config2 1Mx1000")
    b = aligner.batch(ds)
    got, t = aligner.align_plain(b, 448)
    assert (got >= 0).all() and (got % 2 == 0).all()
    pl, tl = ds.lengths()
    bound = 4 * np.minimum(pl, tl) + np.where(pl != tl, 6 + 2 * np.abs(pl.astype(int) - tl), 0)
    assert (got <= bound).all()
    idx = np.random.default_rng(0).choice(n, 2000, replace=False)
    want, _, _ = oracle.wfa_batch([ds.pair(int(i)) for i in idx], 4, 6, 2)
    assert (got[idx] == want).all()
    again, _ = aligner.align_plain(b, 448)
    assert (again == got).all()
    say(f"full config2 kernel_ms={t['kernel_ms']:.2f}")


def test_wfa_gpu_cli_check_mode(repo, tsu, tmp_path):
    """bin/wfa_gpu with check=1 prints the reference's success banner."""
    ds = tsu.Dataset.generate(2000, seed=SEED + 10, len_min=128, err_min=0.02)
    ds.write(str(tmp_path / "sequences.txt"), 2)
    say("cli check 2000x128")
    r = subprocess.run([os.path.join(repo, "bin", "wfa_gpu"), "4", "6", "2", "sequences.txt", "42", "1"],
                       cwd=tmp_path, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    lines = r.stdout.splitlines()
    assert lines[0].startswith("GPU Time: ")
    assert lines[1].startswith("Estimated GCUPS GPU: : ")
    assert lines[2].startswith("CPU Time: ")
    assert lines[3].startswith("Estimated GCUPS SW: ")
    assert "|ALIGNEMENT FINISHED CORRECTLY|" in r.stdout


def test_wfa_gpu_cli_cigar_mode(repo, tsu, oracle, tmp_path):
    """bin/wfa_gpu with TSUNAMI_MODE=cigar and TSUNAMI_CIGARS writes one
    score<TAB>CIGAR line per pair of sequences.txt (300 pairs of 100-3000 bp
    at 1-12%), equal to the oracle's on the pairs as read back from the
    file; the reference's stdout lines stay the same."""
    ds = tsu.Dataset.generate(300, seed=SEED + 67, len_min=100, len_max=3000, err_min=0.01, err_max=0.12)
    ds.write(str(tmp_path / "sequences.txt"), 2)
    env = dict(os.environ, TSUNAMI_MODE="cigar", TSUNAMI_CIGARS=str(tmp_path / "cigars.tsv"))
    say("cli cigar mode 300 pairs")
    r = subprocess.run([os.path.join(repo, "bin", "wfa_gpu"), "4", "6", "2", "sequences.txt", "42", "0"],
                       cwd=tmp_path, env=env, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.splitlines()[0].startswith("GPU Time: ")
    pairs = tsu.Dataset.read(str(tmp_path / "sequences.txt")).pairs()
    want, ops, _ = oracle.wfa_batch(pairs, 4, 6, 2, cigar=True)
    lines = (tmp_path / "cigars.tsv").read_text().splitlines()
    assert len(lines) == len(pairs)
    for i, line in enumerate(lines):
        sc, cig = line.split("\t")
        assert int(sc) == want[i], i
        assert cig == oracle.rle(ops[i]), i


def test_register_and_lds_kernels_agree(aligner, tsu, oracle, fixtures):
    """4/6/2 pairs run through the single-wave register kernel (capacity
    <= 512) and the multi-wave one (capacity > 512) give the oracle's scores;
    narrower register windows either agree or report overflow."""
    cases = [c for c in fixtures if tuple(c["penalties"]) == (4, 6, 2)]
    pairs = [(c["pattern"], c["text"]) for c in cases]
    pairs += tsu.Dataset.generate(400, seed=SEED + 11, len_min=1, len_max=300, err_min=0.0, err_max=0.3).pairs()
    pairs += [("", ""), ("A", ""), ("", "C"), ("ACGT" * 60, "ACGT" * 60), ("A" * 200, "C" * 200)]
    want, _, _ = oracle.wfa_batch(pairs, 4, 6, 2)
    b = aligner.batch(pairs)
    for W in (512, 520, 2000):
        say(f"reg/lds W={W}")
        got, _ = aligner.align_plain(b, W)
        assert (got == want).all(), W
    for W in (64, 128, 256):
        say(f"reg narrow W={W}")
        got, _ = aligner.align_plain(b, W, raise_on_error=False)
        ok = got != tsu.TSU_EKOVERFLOW
        assert ok.sum() > 0
        assert (got[ok] == want[ok]).all(), W
        assert (got[~ok] == tsu.TSU_EKOVERFLOW).all()


def test_multiwave_register_kernel(aligner, tsu, oracle):
    """Capacities of 513..8192 diagonals run on the multi-wave register kernels
    (8-chunk waves up to 1024 diagonals, 16-chunk waves beyond; chunks dealt
    round-robin, edges exchanged through LDS).
    Unequal lengths put I/D cells out of bounds at both sequence ends (the
    trim across waves); narrow windows report overflow, wide ones never."""
    ds = tsu.Dataset.generate(300, seed=SEED + 14, len_min=200, len_max=4000, err_min=0.0, err_max=0.2)
    pairs = ds.pairs()
    rng = np.random.default_rng(14)
    for _ in range(30):
        pairs.append(("".join(rng.choice(list("ACGT"), int(rng.integers(50, 1500)))),
                      "".join(rng.choice(list("ACGT"), int(rng.integers(50, 1500))))))
    pairs.append(("ACGT" * 900, "ACGT" * 900))
    pairs.append(("A" * 100, "C" * 3000))
    want, _, _ = oracle.wfa_batch(pairs, 4, 6, 2)
    b = aligner.batch(pairs)
    for W in (600, 1100, 2100, 4100, 8192):
        say(f"multi-wave W={W}")
        got, _ = aligner.align_plain(b, W, raise_on_error=False)
        ok = got != tsu.TSU_EKOVERFLOW
        assert (got[ok] == want[ok]).all(), W
        assert ok.sum() > len(pairs) // 3, W
        if W == 8192:  # the window covers every diagonal of every pair
            assert ok.all()


def test_wide_windows_lds_ring_kernel(aligner, tsu, oracle):
    """Capacities above 1024 diagonals (up to 16384) run on waves of 16
    chunks whose older M scores sit in an int16 LDS ring (4, 8 or 16 waves).
    Long divergent pairs (12-20 kbp at 12-15%) need 9.5k-14.8k diagonals;
    random and very unequal pairs put I/D cells out of bounds at both ends
    (the waves synchronise with their neighbours only and NULL such cells
    where computed, in score-only and backtrace launches).  Narrow windows report overflow; 16384 covers
    every pair here.  The batched form (initial capacity 16384) agrees."""
    ds = tsu.Dataset.generate(16, seed=SEED + 16, len_min=12000, len_max=20000, err_min=0.12, err_max=0.15)
    pairs = ds.pairs()
    rng = np.random.default_rng(16)
    for _ in range(4):
        pairs.append(("".join(rng.choice(list("ACGT"), int(rng.integers(2000, 6000)))),
                      "".join(rng.choice(list("ACGT"), int(rng.integers(2000, 6000))))))
    pairs.append(("A" * 300, "C" * 7000))
    want, _, _ = oracle.wfa_batch(pairs, 4, 6, 2)
    b = aligner.batch(pairs)
    for W in (3000, 6000, 9000, 16384):
        say(f"wide W={W}")
        got, _ = aligner.align_plain(b, W, raise_on_error=False)
        ok = got != tsu.TSU_EKOVERFLOW
        assert (got[ok] == want[ok]).all(), W
        if W == 16384:
            assert ok.all()
    say("wide batched")
    got, _ = aligner.align_batched(b)
    assert (got == want).all()


def test_wide_windows_doubled_offsets(aligner, tsu, oracle):
    """Sequences of at most 16383 bases run the 16-chunk kernel with doubled
    offsets: 4-16 kbp pairs at 10-15% (plus random, very unequal and
    all-mismatch pairs: I/D out of bounds at both ends) through plain windows
    of 3000-16384 diagonals and the batched form.  An all-mismatch pair of
    8190 bases scores 16380 scaled, just under the variant's score limit."""
    ds = tsu.Dataset.generate(12, seed=SEED + 26, len_min=4000, len_max=16383, err_min=0.10, err_max=0.15)
    pairs = ds.pairs()
    rng = np.random.default_rng(26)
    for _ in range(3):
        pairs.append(("".join(rng.choice(list("ACGT"), int(rng.integers(2000, 6000)))),
                      "".join(rng.choice(list("ACGT"), int(rng.integers(2000, 6000))))))
    pairs += [("A" * 300, "C" * 7000), ("G" * 16383, "G" * 16000 + "T" * 383)]
    want, _, _ = oracle.wfa_batch(pairs, 4, 6, 2)
    b = aligner.batch(pairs)
    for W in (3000, 9000, 16384):
        say(f"doubled W={W}")
        got, _ = aligner.align_plain(b, W, raise_on_error=False)
        ok = got != tsu.TSU_EKOVERFLOW
        assert (got[ok] == want[ok]).all(), W
        if W == 16384:
            assert ok.all()
    say("doubled batched")
    got, _ = aligner.align_batched(b)
    assert (got == want).all()
    mm = [("A" * 8190, "C" * 8190)]
    say("doubled all-mismatch 8190")
    got, _ = aligner.align_plain(aligner.batch(mm), 16383)
    assert int(got[0]) == 4 * 8190
    got, _ = aligner.align_batched(aligner.batch(mm))
    assert int(got[0]) == 4 * 8190


def test_autotune_then_batched(tsu, oracle):
    """The autotuner fills a per-bucket table; batched results stay exact."""
    ds = tsu.Dataset.generate(3000, seed=SEED + 12, len_min=100, len_max=3000, err_min=0.01, err_max=0.15)
    want, _, _ = oracle.wfa_batch(ds.pairs(), 4, 6, 2)
    a = tsu.Aligner(4, 6, 2)
    try:
        b = a.batch(ds)
        say("autotune 3000 mixed")
        table = a.autotune(b, 512)
        assert table and all(e["waves_per_pair"] in (1, 2, 4, 8, 16) for e in table)
        got, _ = a.align_batched(b)
        assert (got == want).all()
        # a saved table restores as it was; a bad entry leaves it unchanged
        a.set_tune_table([])
        assert a.tune_table() == []
        a.set_tune_table(table)
        assert a.tune_table() == table
        for bad in ({"bucket_bound": 1000}, {"bucket_bound": 2048, "waves_per_pair": 3},
                    {"bucket_bound": 2048, "capacity": -1}, {"bucket_bound": 2048, "waves_per_pair": 4, "capacity": 512}):
            with pytest.raises(tsu.TsuError):
                a.set_tune_table([bad])
        with pytest.raises(tsu.TsuError):
            a.set_tune_table([{"bucket_bound": 2048}, {"bucket_bound": 2048}])
        assert a.tune_table() == table
        # a pinned first window of 64 diagonals: pairs re-run wider, same scores
        a.set_tune_table([{"waves_per_pair": 1, "capacity": 64, "bucket_bound": bb} for bb in (512, 1024, 2048, 4096)])
        got, t = a.align_batched(b)
        assert (got == want).all() and t["retried"] > 0
    finally:
        a.close()


def test_long_read_cigars(aligner, tsu, oracle):
    """Config-3 shape with CIGARs: 10,000 bp at 10% (4 pairs, ring in HBM)."""
    ds = tsu.Dataset.generate(4, seed=SEED + 13, len_min=10000, err_min=0.10)
    pairs = ds.pairs()
    want, ops, _ = oracle.wfa_batch(pairs, 4, 6, 2, cigar=True)
    say("cigar long 4x10000")
    got, cig, _ = aligner.align_cigar(aligner.batch(ds))
    assert (got == want).all()
    for i, (p, t) in enumerate(pairs):
        assert cig[i] == oracle.rle(ops[i]), i


def test_wide_cigars_on_register_kernel(aligner, tsu, oracle):
    """CIGARs of pairs whose capacity lands above 1024 diagonals come from the
    16-chunk multi-wave kernel's records (8 and 16 waves here): 12-14 kbp at
    12% and 6 kbp at 15%; same CIGARs as the oracle, which re-score to the
    score."""
    ds = tsu.Dataset.generate(3, seed=SEED + 19, len_min=12000, len_max=14000, err_min=0.12)
    pairs = ds.pairs() + tsu.Dataset.generate(3, seed=SEED + 23, len_min=6000, err_min=0.15).pairs()
    want, ops, _ = oracle.wfa_batch(pairs, 4, 6, 2, cigar=True)
    say("cigar wide 6 pairs")
    got, cig, _ = aligner.align_cigar(aligner.batch(pairs))
    assert (got == want).all()
    for i, (p, t) in enumerate(pairs):
        assert cig[i] == oracle.rle(ops[i]), i
        assert oracle.rescore(oracle.expand(cig[i]), p, t, 4, 6, 2) == got[i]


def test_cigar_single_pass_falls_back(aligner, tsu, oracle):
    """CIGARs are first recorded in one pass per length bucket, without a
    score pass, with records sized for the highest score the bucket's window
    holds.  Unrelated 2,500-3,500 bp sequences (scores far above that) among
    3-3.5 kbp pairs at 5% overflow the window or pass that step bound and
    take the score-first path; every CIGAR equals the oracle's and re-scores
    to the score, and both paths ran record launches."""
    rng = np.random.default_rng(41)
    pairs = tsu.Dataset.generate(6, seed=SEED + 41, len_min=3000, len_max=3500, err_min=0.05).pairs()
    for n in (2500, 3000, 3500):
        pairs.append(("".join(rng.choice(list("ACGT"), n)), "".join(rng.choice(list("ACGT"), n))))
    want, ops, _ = oracle.wfa_batch(pairs, 4, 6, 2, cigar=True)
    say("cigar single pass + fallback, 9 pairs")
    got, cig, t = aligner.align_cigar(aligner.batch(pairs))
    assert (got == want).all(), np.nonzero(got != want)[0]
    for i, (p, t_) in enumerate(pairs):
        assert cig[i] == oracle.rle(ops[i]), i
        assert oracle.rescore(oracle.expand(cig[i]), p, t_, 4, 6, 2) == got[i]
    assert t["record_launches"] >= 2 and t["checkpoint_launches"] == 0


def test_cigars_staged_copy_several_slices(aligner, tsu, oracle):
    """100,000 pairs of 1000 bp at 5%: the run pool (2 * (s / 2) + 2 words
    per pair for scaled score s, about 50 MB) reaches the host in 32 MiB
    staged slices, the next slice's copy overlapping the host copy of the
    current one; 1,500 pairs spread over the batch equal the oracle's
    scores and CIGARs."""
    ds = tsu.Dataset.generate(100000, seed=SEED + 61, len_min=1000, err_min=0.05)
    say("cigar 100000x1000, staged copy")
    got, cig, _ = aligner.align_cigar(aligner.batch(ds))
    pool = int((2 * (got.astype(np.int64) // 2 // 2) + 2).sum()) * 4
    assert pool > 32 << 20, pool
    idx = np.linspace(0, len(ds) - 1, 1500).astype(int)
    pairs = [ds.pair(int(i)) for i in idx]
    want, ops, _ = oracle.wfa_batch(pairs, 4, 6, 2, cigar=True)
    assert (got[idx] == want).all()
    for k, i in enumerate(idx):
        assert cig[i] == oracle.rle(ops[k]), int(i)


@pytest.mark.parametrize("beside", [False, True])
def test_cigar_pairs_wider_than_register_windows(aligner, tsu, oracle, beside):
    """18-19 kbp pairs at 20-22% need about 20,500 diagonals, more than the
    16-chunk kernel's widest window: they overflow its single pass and are
    recorded in one LDS-kernel launch over their whole diagonal range, sized
    by their lengths (no score pass), beside 17-19 kbp pairs at 4% that the
    single pass finishes.  With 3-4 kbp pairs too (beside=True) that launch
    runs on the second stream while their single-pass launch runs.  Same
    scores and CIGARs as the oracle."""
    pairs = (tsu.Dataset.generate(2, seed=SEED + 47, len_min=18000, len_max=20000, err_min=0.20, err_max=0.22).pairs() +
             tsu.Dataset.generate(2, seed=SEED + 53, len_min=17000, len_max=19000, err_min=0.04).pairs())
    if beside:
        pairs += tsu.Dataset.generate(40, seed=SEED + 59, len_min=3000, len_max=4000, err_min=0.02, err_max=0.08).pairs()
    want, ops, _ = oracle.wfa_batch(pairs, 4, 6, 2, cigar=True)
    say(f"cigar wider than the register windows, {len(pairs)} pairs")
    got, cig, t = aligner.align_cigar(aligner.batch(pairs))
    assert (got == want).all(), np.nonzero(got != want)[0]
    for i, (p, t_) in enumerate(pairs):
        assert cig[i] == oracle.rle(ops[i]), i
        assert oracle.rescore(oracle.expand(cig[i]), p, t_, 4, 6, 2) == got[i]
    assert t["record_launches"] >= (3 if beside else 2) and t["checkpoint_launches"] == 0



def test_neighbour_sync_poll_limit_fails_cleanly(tsu, oracle):
    """Score-only launches of 1,025-16,384 diagonals synchronise each wave
    with its two neighbours through bounded polls.  With the bound forced
    down to one poll, waves that would have to wait give up: every pair still
    ends, with its exact score or TSU_EINTERNAL, and the default bound gives
    exact scores again on the same context."""
    ds = tsu.Dataset.generate(6, seed=SEED + 33, len_min=9000, len_max=14000, err_min=0.10, err_max=0.14)
    pairs = ds.pairs()
    want, _, _ = oracle.wfa_batch(pairs, 4, 6, 2)
    a = tsu.Aligner(4, 6, 2)
    try:
        b = a.batch(pairs)
        a.set_poll_limit(1)
        say("poll limit 1, W=16384")
        got, _ = a.align_plain(b, 16384, raise_on_error=False)
        ok = got >= 0
        assert (got[ok] == want[ok]).all()
        assert (got[~ok] == tsu.TSU_EINTERNAL).all()
        a.set_poll_limit(0)
        say("default poll limit, W=16384")
        got, _ = a.align_plain(b, 16384)
        assert (got == want).all()
        with pytest.raises(tsu.TsuError):
            a.set_poll_limit(-1)
    finally:
        a.close()


def test_neighbour_sync_ragged_random_pairs(aligner, tsu, oracle):
    """The neighbour-synchronised 16-chunk launches on 120 ragged pairs: a
    seeded pattern of 1-6 kbp, a text that is an edited copy (0-20%) with a
    random head or tail cut or added, or an unrelated sequence, so cells go
    out of bounds at both ends at many scores.  Plain windows of 3000, 6000
    and 16384 diagonals (4, 8 and 16 waves) and the batched form give the
    oracle's scores wherever no overflow is reported."""
    rng = np.random.default_rng(35)
    alphabet = np.array(list("ACGT"))
    pairs = []
    for i in range(120):
        n = int(rng.integers(1000, 6000))
        p = alphabet[rng.integers(0, 4, n)]
        kind = i % 4
        if kind == 3:
            t = alphabet[rng.integers(0, 4, int(rng.integers(500, 6000)))]
        else:
            t = list(p)
            for _ in range(int(rng.random() * 0.2 * n)):
                j = int(rng.integers(0, len(t)))
                r = int(rng.integers(0, 3))
                if r == 0:
                    t[j] = alphabet[(int(np.nonzero(alphabet == t[j])[0][0]) + 1 + int(rng.integers(0, 3))) % 4]
                elif r == 1:
                    t.insert(j, alphabet[int(rng.integers(0, 4))])
                else:
                    del t[j]
            cut = int(rng.integers(0, n // 3))
            if kind == 1:
                t = t[cut:]
            elif kind == 2:
                t = t + list(alphabet[rng.integers(0, 4, cut)])
        pairs.append(("".join(p), "".join(t)))
    want, _, _ = oracle.wfa_batch(pairs, 4, 6, 2)
    b = aligner.batch(pairs)
    for W in (3000, 6000, 16384):
        say(f"ragged W={W}")
        got, _ = aligner.align_plain(b, W, raise_on_error=False)
        ok = got != tsu.TSU_EKOVERFLOW
        assert ok.sum() > 0
        assert (got[ok] == want[ok]).all(), (W, np.nonzero(got[ok] != want[ok])[0][:5])
    say("ragged batched")
    got, _ = aligner.align_batched(b)
    assert (got == want).all()
