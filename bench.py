"""Benchmark of the gap-affine WFA hot path on one MI355X.

Default workload (the configuration the headline metric is quoted on):
1,000,000 pairs of 1000 bp, text = pattern with 50 mixed edits (substitution /
insertion / deletion, 1/3 each), penalties 4/6/2, plain form (one workgroup
per pair), sequences 2-bit packed and resident in HBM before timing starts.

A step is one pass of the plain kernel over all pairs plus the copy of the
int32 scores to the host.  value = pairs * steps / wall time of the timed
steps (alignments/s).  The kernel's own duration comes from HIP events
recorded on the stream it runs on; the roofline figure divides the
algorithmic HBM bytes of a launch (2-bit sequences read once + 4-byte score
written: 8*(ceil(p/32)+ceil(t/32))+4 per pair) by that duration.

The path's binding roof is integer VALU issue, not HBM (the sequences are
read once, the wavefronts stay on chip).  "roofline_valu" prices the work
with a fixed op model -- 11 int32 lane-ops per cell-triple (one I, D and M
update) and 16 per 32-base extend compare, the counts taken from the oracle
on the CPU sample -- against the chip's int32 VALU peak (256 CUs x 64 lanes
x 2.4 GHz).

The CPU baseline is the oracle's exact WFA (C, OpenMP over pairs, every host
thread), timed on the first pairs of the same workload, and its scores double
as a bit-exactness check of the timed GPU run.

The batched and CIGAR forms run the autotuner on a sample first (untimed);
with --compare-plain the batched form also times the plain form on the same
pairs.

The default run (plain form, 1M x 1000 bp) also measures, each in a child
process after the headline, and adds to the same JSON line:
  batched_cfg4  500,000 pairs, 100-20,000 bp log-uniform, 1-15%: autotuned
                batched form vs the plain form on the same pairs;
  served_cfg5   Poisson arrivals at three rates, requests of U{1..256} pairs
                of 100-10,000 bp at 1-15%: p50 / p99 latency and throughput;
  cigar_cfg3    50,000 pairs of 10,000 bp at 10%: scores and CIGARs.
(--headline-only skips them.)

    python bench.py [--steps K] [--warmup W] [--form plain|batched|cigar|served|refgrid|twopiece]
"""
import argparse
import json
import os
import subprocess
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

SEED = 0x5453554E
HBM_PEAK_GBS = 8000.0
VALU_PEAK_LANE_OPS = 256 * 64 * 2.4e9   # int32 lane-ops/s: CUs x lanes per clock x peak engine clock
OPS_PER_CELL, OPS_PER_EXTEND = 11, 16   # op model: per cell-triple, per 32-base extend compare
PEN = (4, 6, 2)
METRIC = "alignments/sec + equiv. GCUPS at 1 GPU, bit-exact; p50/p99 latency when served"

WORKLOADS = {
    # name: (pairs, len_min, len_max, err_min, err_max, seed offset)
    "cfg1": (10_000, 100, 100, 0.02, 0.02, 1),
    "cfg2": (1_000_000, 1000, 1000, 0.05, 0.05, 2),
    "cfg3": (50_000, 10_000, 10_000, 0.10, 0.10, 3),
    "cfg4": (500_000, 100, 20_000, 0.01, 0.15, 4),
    # long reads for the backtrace-memory comparison (--form cigar)
    "long100k": (256, 100_000, 100_000, 0.05, 0.05, 6),
}


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def algorithmic_bytes(pl, tl):
    return int((8 * ((pl.astype(np.int64) + 31) // 32 + (tl.astype(np.int64) + 31) // 32) + 4).sum())


def host_cpu():
    """The host's CPU model and the cores this process may run on."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"nproc": len(os.sched_getaffinity(0)), "cpu_count": os.cpu_count(), "cpu_model": model}


ORACLE_NOTE = ("the oracle: the C restatement of WFA2-lib's exact gap-affine WFA kept literal for parity "
               "(unscaled penalties, byte-wise extend, per-score wavefront allocation), OpenMP dynamic over "
               "pairs; WFA2-lib itself is absent offline, and no optimised CPU aligner stands in for it")


def cpu_baseline(pairs_fn, n_total, target_cpu_s=20.0):
    """Oracle WFA on the first pairs of the workload, every host thread."""
    from oracle import oracle
    threads = oracle.max_threads()
    n = min(n_total, 2000)
    t0 = time.perf_counter()
    sample = pairs_fn(0, n)
    scores, _, cnt = oracle.wfa_batch(sample, *PEN, threads=threads)
    dt = time.perf_counter() - t0
    # grow the sample to about target_cpu_s of CPU work (threads x wall)
    per_pair = dt * threads / max(n, 1)
    n2 = int(min(n_total, max(n, target_cpu_s / max(per_pair, 1e-9))))
    if n2 > n:
        sample = pairs_fn(0, n2)
        t0 = time.perf_counter()
        scores, _, cnt = oracle.wfa_batch(sample, *PEN, threads=threads)
        dt = time.perf_counter() - t0
        n = n2
    return {"value": n / dt, "unit": "alignments/s", "cores": threads, "kind": "port",
            "sample": f"first {n} pairs of the workload, oracle exact WFA (C, OpenMP, {threads} threads), "
                      f"{dt:.2f} s wall",
            "omp_threads": threads, **host_cpu(), "note": ORACLE_NOTE}, scores, cnt


def read_traffic(name):
    path = os.path.join(REPO, "profiles", f"traffic_{name}.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f).get("bytes_per_launch")


def run_served(a, tsu, args, cpu_sample):
    """Open-loop Poisson arrivals at three rates (fractions of the measured
    batched capacity); each request holds U{1..256} pairs of length
    log-uniform in [100, 10000] bp and 1-15% mixed edits."""
    import threading
    rng = np.random.default_rng(SEED + 5)
    n_req = args.requests
    sizes = rng.integers(1, 257, n_req)
    total = int(sizes.sum())
    log(f"served: generating {n_req} requests, {total} pairs")
    ds = tsu.Dataset.generate(total, seed=SEED + 5, len_min=100, len_max=10_000, err_min=0.01, err_max=0.15)
    firsts = np.concatenate([[0], np.cumsum(sizes)[:-1]])
    # requests are submitted pre-packed (tsu_serve_submit_packed), so the
    # open-loop generator is not limited by packing on its own thread
    words, pairs = ds.packed()
    req_words, req_pairs = [], []
    for i in range(n_req):
        pr = pairs[firsts[i]:firsts[i] + sizes[i]].copy()
        w0 = int(pr["p_off"][0])
        w1 = int(pr["t_off"][-1] + (pr["t_len"][-1] + 15) // 16)
        pr["p_off"] -= w0
        pr["t_off"] -= w0
        req_words.append(words[w0:w1])
        req_pairs.append(pr)
    # capacity: batched form over all pairs of the stream
    b = a.batch(ds)
    ref, t = a.align_batched(b)
    ref, t = a.align_batched(b)
    cap_pairs = total / (t["wall_ms"] / 1e3)
    plain = None
    if args.compare_plain:
        # the plain form on the same pairs: one launch, capacity for the
        # score bound of the longest, most divergent request pair
        x, o, e = PEN
        W = 2 * ((int(round(0.15 * 10_000)) * (o + e) - o) // e) + 3
        sc_p, tp = a.align_plain(b, W)
        sc_p, tp = a.align_plain(b, W)
        plain = {"alignments_per_s": total / (tp["wall_ms"] / 1e3), "kernel_ms": tp["kernel_ms"], "wf_length": W,
                 "same_scores": bool((sc_p == ref).all())}
        log(f"served: plain form on the same pairs {plain}")
    b.close()
    log(f"served: batched capacity {cap_pairs:.0f} pairs/s")
    out = []
    for frac in (0.3, 0.6, 0.9):
        rate_req = frac * cap_pairs / float(sizes.mean())
        gaps = rng.exponential(1.0 / rate_req, n_req)
        srv = tsu.Server(a, hbm_budget_bytes=args.serve_budget << 20, max_inflight=args.inflight,
                         max_batch_pairs=args.batch_pairs)
        tickets = [None] * n_req
        ready = [threading.Event() for _ in range(n_req)]

        def submitter():
            t0 = time.perf_counter()
            due = 0.0
            for i in range(n_req):
                due += gaps[i]
                d = t0 + due - time.perf_counter()
                if d > 0:
                    time.sleep(d)
                tickets[i] = srv.submit_packed(req_words[i], req_pairs[i])
                ready[i].set()

        th = threading.Thread(target=submitter)
        th.start()
        lat, sub, done, qwait, exact = [], [], [], [], True
        for i in range(n_req):
            ready[i].wait()
            st, sc, stats = srv.wait(tickets[i], int(sizes[i]))
            if st != tsu.TSU_OK:
                raise RuntimeError(f"request {i}: status {st}")
            f = int(firsts[i])
            exact &= bool((sc == ref[f:f + int(sizes[i])]).all())
            lat.append((stats["done_ns"] - stats["submit_ns"]) / 1e6)
            qwait.append((stats["admit_ns"] - stats["submit_ns"]) / 1e6)
            sub.append(stats["submit_ns"]); done.append(stats["done_ns"])
        th.join()
        srv.close()
        span = (max(done) - min(sub)) / 1e9
        res = {"offered_fraction": frac, "offered_requests_per_s": rate_req, "requests": n_req, "pairs": total,
               "throughput_pairs_per_s": total / span, "p50_ms": float(np.percentile(lat, 50)),
               "p99_ms": float(np.percentile(lat, 99)), "mean_ms": float(np.mean(lat)),
               "mean_queue_ms": float(np.mean(qwait)),
               "same_as_batched": exact}
        log(f"served {frac:.1f}: {res}")
        out.append(res)
    res = {"capacity_pairs_per_s": cap_pairs, "rates": out}
    if plain:
        plain["served_speedup_at_0.9"] = out[-1]["throughput_pairs_per_s"] / plain["alignments_per_s"]
        res["plain_same_pairs"] = plain
    return res


# The reference's own benchmark grid (run_h100-a100.sh): generate_dataset N L E
# (substitution-only edits at E%) then wfa_gpu 4 6 2 sequences.txt W 1.
REF_GRID = [
    (500_000, 128, 1, 21), (500_000, 512, 1, 21), (50_000, 1024, 1, 200), (10_000, 10240, 1, 350),
    (500_000, 128, 2, 42), (500_000, 512, 2, 42), (50_000, 1024, 2, 250), (10_000, 10240, 2, 850),
    (500_000, 128, 5, 105), (500_000, 512, 5, 105), (50_000, 1024, 5, 250), (10_000, 10240, 5, 2200),
]
# Peak equivalent GCUPS the reference reports for its best configuration
# (README.md:6; the GPU is not named there).
REF_PEAK_GCUPS = 42604.98


def run_refgrid(a, tsu, args):
    """Plain form on the reference's own grid of configurations: equivalent
    GCUPS (sum of p*t over the kernel time, wfa_gpu.cu:130-133).  The
    capacity starts at the reference's wf_length and doubles while a pair
    overflows (the reference writes past its ring instead)."""
    out = []
    for n, L, pct, w_ref in REF_GRID:
        if args.pairs:
            n = min(n, args.pairs)
        ds = tsu.Dataset.generate(n, seed=SEED + 100 + 10 * pct + L, len_min=L, err_min=pct / 100.0, edit_mode=1)
        b = a.batch(ds)
        W = w_ref
        while True:
            sc, _ = a.align_plain(b, W, raise_on_error=False)
            if not (sc == tsu.TSU_EKOVERFLOW).any():
                break
            W *= 2
        kms = []
        for _ in range(max(args.steps, 1)):
            sc, t = a.align_plain(b, W)
            kms.append(t["kernel_ms"])
        kms = float(np.median(kms))
        res = {"pairs": n, "length": L, "error_pct": pct, "edits": "substitutions", "wf_length_reference": w_ref,
               "wf_length": W, "kernel_ms": kms, "alignments_per_s": n / (kms / 1e3),
               "gcups": n * L * L / (kms / 1e3) / 1e9}
        if not args.no_cpu:
            m = min(n, 20_000 if L <= 1024 else 400)
            cpu, want, _ = cpu_baseline(lambda i0, k: [ds.pair(i) for i in range(i0, i0 + k)], m, target_cpu_s=2.0)
            res["cpu_baseline"] = cpu
            res["bit_exact_sample"] = bool((want == sc[:len(want)]).all())
        log(f"refgrid {L} bp {pct}%: {res}")
        out.append(res)
        b.close()
        ds.close()
    best = max(out, key=lambda r: r["gcups"])
    return {"configs": out, "best_gcups": best["gcups"], "best_config": f"{best['length']} bp {best['error_pct']}%",
            "reference_published_peak_gcups": REF_PEAK_GCUPS,
            "vs_reference_peak": best["gcups"] / REF_PEAK_GCUPS}


def oracle_sample(ds, m, cigar=False):
    """Oracle scores (and op strings) of the first m pairs of ds, to check the GPU's."""
    from oracle import oracle
    sample = [ds.pair(i) for i in range(min(m, len(ds)))]
    scores, ops, _ = oracle.wfa_batch(sample, *PEN, threads=oracle.max_threads(), cigar=cigar)
    return scores, ops


def section_batched_cfg4(tsu, args):
    """The mixed-length workload: autotuned batched form against the plain
    form on the same resident pairs (one capacity for all, sized for the
    score bound of the longest, most divergent pair)."""
    n, lmin, lmax, emin, emax, soff = WORKLOADS["cfg4"]
    if args.pairs:
        n = min(n, args.pairs)
    log(f"batched_cfg4: generating {n} pairs")
    ds = tsu.Dataset.generate(n, seed=SEED + soff, len_min=lmin, len_max=lmax, err_min=emin, err_max=emax)
    pl, tl = ds.lengths()
    cells = float((pl.astype(np.float64) * tl).sum())
    a = tsu.Aligner(*PEN)
    b = a.batch(ds)
    t0 = time.perf_counter()
    tune = a.autotune(b, args.tune_sample)
    tune_s = time.perf_counter() - t0
    log(f"batched_cfg4: autotune {tune_s:.1f} s")
    scores, _ = a.align_batched(b)                 # warmup
    walls, kms = [], []
    for i in range(2):
        t0 = time.perf_counter()
        scores, t = a.align_batched(b)
        walls.append(time.perf_counter() - t0)
        kms.append(t["kernel_ms"])
        log(f"batched_cfg4: batched step {i}: {walls[-1] * 1e3:.1f} ms")
    x, o, e = PEN
    W = 2 * ((int(round(emax * lmax)) * (o + e) - o) // e) + 3
    a.align_plain(b, W)                             # warmup
    t0 = time.perf_counter()
    ps, pt = a.align_plain(b, W)
    pwall = time.perf_counter() - t0
    log(f"batched_cfg4: plain step: {pwall * 1e3:.1f} ms")
    bw = float(np.median(walls))
    res = {"workload": f"cfg4: {n} pairs, length {lmin}-{lmax} log-uniform bp, {emin * 100:g}-{emax * 100:g}% mixed "
                       "edits, penalties 4/6/2", "pairs": n, "seed": SEED + soff,
           "batched": {"alignments_per_s": n / bw, "wall_ms": bw * 1e3, "kernel_ms": float(np.median(kms)),
                       "gcups": cells / bw / 1e9, "launches": t["launches"], "pairs_widened": t["retried"],
                       "steps": len(walls), "autotune_s": tune_s},
           "plain_same_pairs": {"alignments_per_s": n / pwall, "wall_ms": pwall * 1e3, "kernel_ms": pt["kernel_ms"],
                                "wf_length": W},
           "speedup_batched_over_plain": pwall / bw,
           "same_scores": bool((ps == scores).all()),
           "tune": tune}
    if not args.no_cpu:
        want, _ = oracle_sample(ds, 2000)
        res["bit_exact_sample"] = bool((scores[:len(want)] == want).all())
        res["sample_pairs"] = int(len(want))
    b.close()
    a.close()
    return res


def section_served_cfg5(tsu, args):
    """Poisson arrivals at three open-loop rates (run_served)."""
    ns = argparse.Namespace(requests=args.requests, compare_plain=True, serve_budget=args.serve_budget,
                            inflight=args.inflight, batch_pairs=args.batch_pairs)
    a = tsu.Aligner(*PEN)
    res = run_served(a, tsu, ns, None)
    a.close()
    return {"workload": "cfg5: Poisson arrivals at 0.3/0.6/0.9 of the batched capacity, requests of U{1..256} "
                        "pairs, length 100-10000 bp log-uniform, 1-15% mixed edits, penalties 4/6/2",
            "server": {"hbm_budget_mib": args.serve_budget, "max_inflight": args.inflight,
                       "max_batch_pairs": args.batch_pairs}, **res}


def section_cigar_cfg3(tsu, args):
    """Long erroneous reads with CIGARs, at the configured 50,000 pairs."""
    n, lmin, lmax, emin, emax, soff = WORKLOADS["cfg3"]
    if args.pairs:
        n = min(n, args.pairs)
    log(f"cigar_cfg3: generating {n} pairs")
    ds = tsu.Dataset.generate(n, seed=SEED + soff, len_min=lmin, len_max=lmax, err_min=emin, err_max=emax)
    pl, tl = ds.lengths()
    cells = float((pl.astype(np.float64) * tl).sum())
    a = tsu.Aligner(*PEN)
    b = a.batch(ds)
    tune = a.autotune(b, args.tune_sample)
    _, _, tw = a.align_cigar(b, strings=False, count_runs=True)   # warmup; the same CIGARs as the timed call
    t0 = time.perf_counter()
    m = 0 if args.no_cpu else 200
    scores, cigars, t = a.align_cigar(b, strings=m)  # CIGAR strings of the first m pairs only
    wall = time.perf_counter() - t0
    log(f"cigar_cfg3: {wall * 1e3:.1f} ms")
    res = {"workload": f"cfg3: {n} pairs, {lmin} bp, {emin * 100:g}% mixed edits, penalties 4/6/2, CIGAR form",
           "pairs": n, "seed": SEED + soff, "alignments_per_s": n / wall, "wall_ms": wall * 1e3,
           "kernel_ms": t["kernel_ms"], "gcups": cells / wall / 1e9, "launches": t["launches"],
           "backtrace": {"state_bytes": t["state_bytes"], "record_launches": t["record_launches"],
                         "checkpoint_launches": t["checkpoint_launches"], "kernels": sorted(t["kernels"])},
           "tune": tune}
    # HBM: packed sequences and scores, plus the CIGAR runs (4 bytes each) and
    # a run count per pair; traffic from the committed counter runs of this call
    algo = algorithmic_bytes(pl, tl) + 4 * tw["cigar_runs"] + 4 * n
    ach = algo / (t["kernel_ms"] / 1e3) / 1e9
    traffic = read_traffic("wfa_cigar_cfg3") if n == WORKLOADS["cfg3"][0] else None
    res["roofline_hbm"] = {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                           "frac": ach / HBM_PEAK_GBS, "algorithmic_bytes": algo, "cigar_runs": tw["cigar_runs"],
                           "traffic": traffic,
                           "traffic_frac": None if traffic is None else traffic / (t["kernel_ms"] / 1e3) / 1e9 / HBM_PEAK_GBS}
    if m:
        from oracle import oracle
        want, ops = oracle_sample(ds, m, cigar=True)
        res["bit_exact_sample"] = bool((scores[:m] == want).all()) and cigars == [oracle.rle(o) for o in ops]
        res["sample_pairs"] = m
    b.close()
    a.close()
    return res


SECTIONS = {"batched_cfg4": section_batched_cfg4, "served_cfg5": section_served_cfg5,
            "cigar_cfg3": section_cigar_cfg3}


def run_section_child(name, args):
    """One section in a child process (its own device context and memory);
    its progress goes to stderr, its result is its last stdout line."""
    cmd = [sys.executable, os.path.abspath(__file__), "--section", name, "--tune-sample", str(args.tune_sample),
           "--requests", str(args.requests)]
    if args.pairs:
        cmd += ["--pairs", str(args.pairs)]
    if args.no_cpu:
        cmd.append("--no-cpu")
    log(f"section {name}")
    try:
        r = subprocess.run(cmd, stdout=subprocess.PIPE, text=True, timeout=900)
    except subprocess.TimeoutExpired:
        return {"error": "timed out after 900 s"}
    lines = r.stdout.strip().splitlines()
    if r.returncode != 0 or not lines:
        return {"error": f"exit status {r.returncode}"}
    return json.loads(lines[-1])


PEN_2P = (4, 6, 2, 24, 1)  # gap-affine 2-piece: gap of length L costs min(6 + 2L, 24 + L)


def run_twopiece(tsu, args):
    """Plain form under gap-affine 2-piece penalties on pairs of the config-2
    shape (1000 bp, 5% mixed edits); capacity from the score bound (each edit
    costs at most o1 + e1, the range grows by one diagonal per e2)."""
    n = args.pairs or 100_000
    x, o1, e1, o2, e2 = PEN_2P
    ds = tsu.Dataset.generate(n, seed=SEED + 22, len_min=1000, err_min=0.05)
    s_bound = int(round(0.05 * 1000)) * max(x, o1 + e1)
    W = 2 * max((s_bound - o1) // e1, (s_bound - o2) // e2) + 3
    a = tsu.Aligner(x, o1, e1, gap_opening2=o2, gap_extension2=e2)
    b = a.batch(ds)
    sc, t = a.align_plain(b, W)
    kms = []
    for _ in range(max(args.steps, 1)):
        sc, t = a.align_plain(b, W)
        kms.append(t["kernel_ms"])
    kms = float(np.median(kms))
    res = {"metric": "alignments/sec, gap-affine 2-piece, plain form, 1 GPU", "form": "twopiece",
           "penalties": list(PEN_2P), "pairs": n, "length": 1000, "error": 0.05, "wf_length": W,
           "kernel_ms": kms, "alignments_per_s": n / (kms / 1e3), "gcups": n * 1e6 / (kms / 1e3) / 1e9,
           "device": a.device_info()["name"]}
    if not args.no_cpu:
        from oracle import oracle
        m = min(n, 4000)
        sample = [ds.pair(i) for i in range(m)]
        threads = oracle.max_threads()
        t0 = time.perf_counter()
        want, _ = oracle.wfa2p_batch(sample, *PEN_2P, threads=threads)
        dt = time.perf_counter() - t0
        res["cpu_baseline"] = {"value": m / dt, "unit": "alignments/s", "cores": threads, "kind": "port",
                               "sample": f"first {m} pairs, 2-piece oracle WFA (C, OpenMP, {threads} threads), {dt:.2f} s wall"}
        res["bit_exact_sample"] = bool((want == sc[:m]).all())
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--form", default="plain", choices=["plain", "batched", "cigar", "served", "refgrid", "twopiece"])
    ap.add_argument("--workload", default="cfg2", choices=sorted(WORKLOADS))
    ap.add_argument("--pairs", type=int, default=0, help="override the pair count (testing only)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-autotune", action="store_true", help="batched, cigar: use the default shapes")
    ap.add_argument("--tune-sample", type=int, default=4096, help="batched: autotuner sample per bucket")
    ap.add_argument("--compare-plain", action="store_true", help="batched, served: also time the plain form")
    ap.add_argument("--cigar-memory", default="auto", choices=["auto", "records", "checkpoints"],
                    help="cigar: backtrace state (records for every score, or ring checkpoints)")
    ap.add_argument("--checkpoint-every", type=int, default=0, help="cigar: scores between checkpoints, 0 = automatic")
    ap.add_argument("--requests", type=int, default=3000, help="served: requests per rate")
    ap.add_argument("--serve-budget", type=int, default=16384, help="served: HBM budget in MiB")
    ap.add_argument("--inflight", type=int, default=4, help="served: launches in flight")
    ap.add_argument("--batch-pairs", type=int, default=1024, help="served: pairs merged per launch")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="BYTES",
                    help="only run the streaming copy kernel over BYTES (HBM counter calibration)")
    ap.add_argument("--headline-only", action="store_true",
                    help="default form: skip the batched_cfg4, served_cfg5 and cigar_cfg3 sections")
    ap.add_argument("--section", choices=sorted(SECTIONS), help=argparse.SUPPRESS)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              "(pairs are independent; a multi-GPU run would shard them by estimated work)", file=sys.stderr)
        sys.exit(2)

    import tsunami_pact_amd as tsu

    if args.section:
        print(json.dumps(SECTIONS[args.section](tsu, args)), flush=True)
        return
    if args.calibrate_copy:
        a = tsu.Aligner(*PEN)
        ms = a.calibrate_copy(args.calibrate_copy)
        print(json.dumps({"copy_bytes": args.calibrate_copy, "kernel_ms": ms,
                          "GB_per_s": 2 * args.calibrate_copy / (ms / 1e3) / 1e9}), flush=True)
        return
    if args.form == "twopiece":
        print(json.dumps(run_twopiece(tsu, args)), flush=True)
        return
    if args.form == "refgrid":
        a = tsu.Aligner(*PEN)
        res = run_refgrid(a, tsu, args)
        print(json.dumps({"metric": "equiv. GCUPS on the reference's benchmark grid (plain form, 1 GPU)",
                          "form": "refgrid", "device": a.device_info()["name"], **res}), flush=True)
        return
    if args.form == "served":
        a = tsu.Aligner(*PEN)
        res = run_served(a, tsu, args, None)
        print(json.dumps({"metric": METRIC, "form": "served", "workload": "cfg5: Poisson arrivals, requests of "
                          "U{1..256} pairs, length log-uniform 100-10000 bp, 1-15% mixed edits, penalties 4/6/2",
                          "device": a.device_info()["name"], **res}), flush=True)
        return
    n, lmin, lmax, emin, emax, soff = WORKLOADS[args.workload]
    if args.pairs:
        n = args.pairs
    log(f"generating {n} pairs ({args.workload})")
    ds = tsu.Dataset.generate(n, seed=SEED + soff, len_min=lmin, len_max=lmax, err_min=emin, err_max=emax)
    pl, tl = ds.lengths()
    x, o, e = PEN
    # capacity from the score bound: <= round(rate*L) edits of cost <= o+e
    edits = int(round(emax * lmax))
    s_bound = edits * (o + e)
    W = 2 * ((s_bound - o) // e) + 3

    a = tsu.Aligner(*PEN)
    info = a.device_info()
    log(f"device {info}")
    b = a.batch(ds)
    log(f"batch resident: {b.nbytes / 1e6:.1f} MB")

    if args.form == "cigar":
        mode = {"auto": a.CIGAR_AUTO, "records": a.CIGAR_RECORDS, "checkpoints": a.CIGAR_CHECKPOINTS}[args.cigar_memory]
        a.set_cigar_memory(mode, args.checkpoint_every)
    tune = None
    # the CIGAR form sizes its record launches from the batched form's per-bucket
    # capacities (and its score-first pairs run the batched form): same table
    if args.form in ("batched", "cigar") and not args.no_autotune:
        t0 = time.perf_counter()
        tune = a.autotune(b, args.tune_sample)
        log(f"autotune ({time.perf_counter() - t0:.1f} s): {tune}")

    out = np.ones(max(n, 1), np.int32)  # the scores' host array, its pages touched once here

    def step():
        if args.form == "plain":
            return a.align_plain(b, W, out=out)
        if args.form == "batched":
            return a.align_batched(b, out=out)
        s, _, t = a.align_cigar(b, strings=False)  # runs stay in the library's host buffers
        return s, t

    for i in range(args.warmup):
        scores, _ = step()
        log(f"warmup {i} done")
    a.sync()
    kms = []
    t0 = time.perf_counter()
    said = t0
    for i in range(args.steps):
        scores, t = step()
        kms.append(t["kernel_ms"])
        # a line per step for the first steps and then every few seconds
        # (steps of config 1 take tens of microseconds)
        now = time.perf_counter()
        if i < 2 or i == args.steps - 1 or now - said > 5:
            log(f"step {i}: kernel {t['kernel_ms']:.3f} ms")
            said = now
    a.sync()
    wall = time.perf_counter() - t0
    scores = scores.copy()

    value = n * args.steps / wall
    cells = float((pl.astype(np.float64) * tl).sum())
    kernel_ms = float(np.mean(kms))
    algo_bytes = algorithmic_bytes(pl, tl)
    achieved = algo_bytes / (kernel_ms / 1e3) / 1e9
    traffic = read_traffic(f"wfa_{args.form}_{args.workload}")

    plain = None
    if args.form == "batched" and args.compare_plain:
        log(f"plain form on the same pairs (W={W})")
        t0 = time.perf_counter()
        ps, pt = a.align_plain(b, W)
        a.sync()
        pwall = time.perf_counter() - t0
        plain = {"alignments_per_s": n / pwall, "kernel_ms": pt["kernel_ms"], "wf_length": W,
                 "batched_speedup": (n * args.steps / wall) / (n / pwall),
                 "same_scores": bool((ps == scores).all())}
        log(f"plain: {plain}")

    cpu = None
    exact = None
    valu = None
    if not args.no_cpu:
        log("cpu baseline")
        cpu, cpu_scores, cnt = cpu_baseline(lambda s, m: [ds.pair(i) for i in range(s, s + m)], n)
        m = len(cpu_scores)
        exact = bool((scores[:m] == cpu_scores).all())
        if args.form != "cigar":
            ops_per_pair = (OPS_PER_CELL * cnt["cells"] + OPS_PER_EXTEND * cnt["extend_words"]) / m
            ach = ops_per_pair * n / (kernel_ms / 1e3)
            valu = {"bound": "valu", "achieved": ach / 1e12, "peak": VALU_PEAK_LANE_OPS / 1e12,
                    "unit": "TOP/s (int32 lane-ops)", "frac": ach / VALU_PEAK_LANE_OPS,
                    "ops_per_pair": ops_per_pair,
                    "model": f"{OPS_PER_CELL}/cell-triple + {OPS_PER_EXTEND}/extend compare, oracle counts "
                             f"over the CPU sample ({m} pairs)"}

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "alignments/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic",
        "config": {"workload": f"{args.workload}: {n} pairs, length "
                               f"{lmin if lmin == lmax else f'{lmin}-{lmax} log-uniform'} bp, "
                               f"{emin * 100:g}{'' if emin == emax else f'-{emax * 100:g}'}% mixed edits, "
                               f"penalties 4/6/2, {args.form} form",
                   "pairs": n, "form": args.form, "wf_length": W, "penalties": list(PEN), "seed": SEED + soff},
        "gcups": cells / (wall / args.steps) / 1e9,
        "kernel_ms": kernel_ms,
        "launches_per_step": t.get("launches"),
        "pairs_widened_per_step": t.get("retried"),
        "bit_exact_sample": exact,
        "cpu_baseline": cpu,
        "device": info["name"],
    }
    # the binding roof is integer VALU issue; HBM is reported beside it
    hbm = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": achieved / HBM_PEAK_GBS, "traffic": traffic}
    if valu is not None:
        out["roofline"] = {**valu, "traffic": traffic}
        out["roofline_hbm"] = hbm
    else:
        out["roofline"] = hbm
    if args.form == "cigar":
        out["config"]["cigar_memory"] = args.cigar_memory
        out["backtrace"] = {"state_bytes": t.get("state_bytes"), "record_launches": t.get("record_launches"),
                            "checkpoint_launches": t.get("checkpoint_launches")}
    if tune is not None:
        out["tune"] = tune
    if plain is not None:
        out["plain_same_pairs"] = plain
    if args.form == "plain" and args.workload == "cfg2" and not args.headline_only:
        b.close()
        ds.close()
        for name in ("batched_cfg4", "served_cfg5", "cigar_cfg3"):
            out[name] = run_section_child(name, args)
            log(f"{name}: {json.dumps(out[name])[:400]}")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
