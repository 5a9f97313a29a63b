"""Python binding of libtsunami_hip.so (gap-affine WFA on MI355X).

A thin ctypes layer over the C-ABI declared in include/tsunami_wfa.h.  It is
used by the tests and bench.py; the product itself is the shared library and
the two CLIs under bin/ (drop-ins for the reference's wfa_gpu and
generate_dataset, wfa_gpu.cu:43-188, generate_dataset.cpp:11-75).

Nothing here computes alignments: every call goes to the HIP library, and
importing fails loudly when the library has not been built.
"""
from __future__ import annotations

import ctypes
import os
import weakref
from typing import Iterable, List, Optional, Sequence, Tuple

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libtsunami_hip.so")
REPO = os.path.dirname(_HERE)
BIN_DIR = os.path.join(REPO, "bin")

TSU_OK = 0
TSU_EINVAL = -1
TSU_EBADINPUT = -2
TSU_EKOVERFLOW = -3
TSU_ENOMEM = -4
TSU_EHIP = -5
TSU_EBUDGET = -6
TSU_EAGAIN = -7
TSU_EIO = -8
TSU_EINTERNAL = -9
TSU_ECIGAR = -10

OPS = "MXID"

# TSU_KERNEL_* bits (tsu_cigars_kernels): which kernels recorded backtraces
KERNEL_BITS = {0x01: "reg1", 0x02: "ml4", 0x04: "ml8", 0x08: "ml16", 0x10: "lds", 0x20: "lds_ck", 0x40: "2p", 0x80: "2p_reg"}

# Every symbol include/tsunami_wfa.h declares.
EXPORTS = (
    "tsu_create", "tsu_create_2p", "tsu_destroy", "tsu_last_error", "tsu_strerror", "tsu_device_info",
    "tsu_packed_words", "tsu_pack",
    "tsu_dataset_read", "tsu_dataset_write", "tsu_dataset_generate", "tsu_dataset_from_strings",
    "tsu_dataset_size", "tsu_dataset_pair", "tsu_dataset_lengths", "tsu_dataset_header", "tsu_dataset_free",
    "tsu_batch_from_dataset", "tsu_batch_create", "tsu_batch_size", "tsu_batch_device", "tsu_batch_bytes",
    "tsu_batch_free",
    "tsu_align_plain", "tsu_align_fixed", "tsu_align_batched", "tsu_align_batched_noalloc", "tsu_autotune",
    "tsu_tune_table", "tsu_set_tune_table",
    "tsu_align_cigar", "tsu_cigars_get", "tsu_cigars_string", "tsu_cigars_free",
    "tsu_set_cigar_memory", "tsu_cigars_stats", "tsu_cigars_kernels", "tsu_set_poll_limit",
    "tsu_serve_open", "tsu_serve_submit", "tsu_serve_submit_packed", "tsu_serve_poll", "tsu_serve_wait",
    "tsu_serve_close",
    "tsu_calibrate_copy", "tsu_sync",
)


class TsuError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"{msg} (status {status})")
        self.status = status


class Pair(ctypes.Structure):
    _fields_ = [("p_off", ctypes.c_int64), ("t_off", ctypes.c_int64),
                ("p_len", ctypes.c_int32), ("t_len", ctypes.c_int32)]


# numpy layout of tsu_pair (24 bytes, naturally aligned)
PAIR_DTYPE = np.dtype([("p_off", np.int64), ("t_off", np.int64), ("p_len", np.int32), ("t_len", np.int32)])


class Timing(ctypes.Structure):
    _fields_ = [("kernel_ms", ctypes.c_double), ("wall_ms", ctypes.c_double),
                ("h2d_ms", ctypes.c_double), ("d2h_ms", ctypes.c_double),
                ("pairs", ctypes.c_int64), ("cells", ctypes.c_double),
                ("launches", ctypes.c_int32), ("retried", ctypes.c_int32)]

    def as_dict(self) -> dict:
        return {f: getattr(self, f) for f, _ in self._fields_}


class GenCfg(ctypes.Structure):
    _fields_ = [("seed", ctypes.c_uint64), ("len_min", ctypes.c_int32), ("len_max", ctypes.c_int32),
                ("err_min", ctypes.c_double), ("err_max", ctypes.c_double),
                ("edit_mode", ctypes.c_int32), ("reserved", ctypes.c_int32)]


class TuneEntry(ctypes.Structure):
    _fields_ = [("waves_per_pair", ctypes.c_int32), ("capacity", ctypes.c_int32),
                ("bucket_bound", ctypes.c_int32), ("reserved", ctypes.c_int32)]


class ServeCfg(ctypes.Structure):
    _fields_ = [("hbm_budget_bytes", ctypes.c_int64), ("max_inflight", ctypes.c_int32),
                ("max_batch_pairs", ctypes.c_int32)]


class ReqStats(ctypes.Structure):
    _fields_ = [("submit_ns", ctypes.c_int64), ("admit_ns", ctypes.c_int64), ("done_ns", ctypes.c_int64),
                ("status", ctypes.c_int32), ("n_pairs", ctypes.c_int32)]


_lib = None


def library() -> ctypes.CDLL:
    """Loads the in-tree libtsunami_hip.so; raises if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is not built: run `make -C tsunami_pact_amd` "
                          "(or __graft_entry__.build())")
    lib = ctypes.CDLL(LIB_PATH)
    vp, i32, i64 = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64
    P = ctypes.POINTER
    sig = {
        "tsu_create": (ctypes.c_int, [P(vp), ctypes.c_int, i32, i32, i32]),
        "tsu_create_2p": (ctypes.c_int, [P(vp), ctypes.c_int, i32, i32, i32, i32, i32]),
        "tsu_destroy": (None, [vp]),
        "tsu_last_error": (ctypes.c_char_p, [vp]),
        "tsu_strerror": (ctypes.c_char_p, [ctypes.c_int]),
        "tsu_device_info": (ctypes.c_int, [vp, P(i32), P(i32), ctypes.c_char_p, i32]),
        "tsu_packed_words": (i64, [i32]),
        "tsu_pack": (ctypes.c_int, [ctypes.c_char_p, i32, P(ctypes.c_uint32)]),
        "tsu_dataset_read": (ctypes.c_int, [ctypes.c_char_p, P(vp), ctypes.c_char_p, i32]),
        "tsu_dataset_write": (ctypes.c_int, [vp, ctypes.c_char_p, i32]),
        "tsu_dataset_generate": (ctypes.c_int, [P(GenCfg), i64, P(vp)]),
        "tsu_dataset_from_strings": (ctypes.c_int, [P(ctypes.c_char_p), P(i32), P(ctypes.c_char_p), P(i32), i64, P(vp)]),
        "tsu_dataset_size": (i64, [vp]),
        "tsu_dataset_pair": (ctypes.c_int, [vp, i64, P(ctypes.c_void_p), P(i32), P(ctypes.c_void_p), P(i32)]),
        "tsu_dataset_lengths": (ctypes.c_int, [vp, P(i32), P(i32)]),
        "tsu_dataset_header": (ctypes.c_int, [vp, P(i32), P(i32), P(i32)]),
        "tsu_dataset_free": (None, [vp]),
        "tsu_batch_from_dataset": (ctypes.c_int, [vp, vp, i64, i64, P(vp)]),
        "tsu_batch_create": (ctypes.c_int, [vp, P(ctypes.c_uint32), i64, P(Pair), i32, P(vp)]),
        "tsu_batch_size": (i32, [vp]),
        "tsu_batch_device": (ctypes.c_int, [vp, P(vp), P(i64), P(vp), P(vp)]),
        "tsu_batch_bytes": (i64, [vp]),
        "tsu_batch_free": (None, [vp]),
        "tsu_align_plain": (ctypes.c_int, [vp, vp, i32, P(i32), P(Timing)]),
        "tsu_align_fixed": (ctypes.c_int, [vp, ctypes.c_char_p, ctypes.c_char_p, i32, i32, i32, P(i32), P(Timing)]),
        "tsu_align_batched": (ctypes.c_int, [vp, vp, P(i32), P(Timing)]),
        "tsu_align_batched_noalloc": (ctypes.c_int, [vp, vp, i64, vp, P(Pair), i32, vp, P(i32), P(Timing)]),
        "tsu_autotune": (ctypes.c_int, [vp, vp, i32]),
        "tsu_tune_table": (ctypes.c_int, [vp, P(TuneEntry), i32, P(i32)]),
        "tsu_set_tune_table": (ctypes.c_int, [vp, P(TuneEntry), i32]),
        "tsu_align_cigar": (ctypes.c_int, [vp, vp, P(i32), P(vp), P(Timing)]),
        "tsu_cigars_get": (ctypes.c_int, [vp, i64, P(P(ctypes.c_uint32)), P(i32)]),
        "tsu_cigars_string": (ctypes.c_int, [vp, i64, ctypes.c_char_p, i32]),
        "tsu_cigars_free": (None, [vp]),
        "tsu_set_cigar_memory": (ctypes.c_int, [vp, i32, i32, i64]),
        "tsu_cigars_stats": (ctypes.c_int, [vp, P(i64), P(i32), P(i32)]),
        "tsu_cigars_kernels": (ctypes.c_int, [vp, P(ctypes.c_uint32)]),
        "tsu_set_poll_limit": (ctypes.c_int, [vp, i32]),
        "tsu_serve_open": (ctypes.c_int, [vp, P(ServeCfg), P(vp)]),
        "tsu_serve_submit": (ctypes.c_int, [vp, P(ctypes.c_char_p), P(i32), P(ctypes.c_char_p), P(i32), i32, P(ctypes.c_uint64)]),
        "tsu_serve_submit_packed": (ctypes.c_int, [vp, P(ctypes.c_uint32), i64, P(Pair), i32, P(ctypes.c_uint64)]),
        "tsu_serve_poll": (ctypes.c_int, [vp, ctypes.c_uint64, P(i32), P(ReqStats)]),
        "tsu_serve_wait": (ctypes.c_int, [vp, ctypes.c_uint64, P(i32), P(ReqStats)]),
        "tsu_serve_close": (ctypes.c_int, [vp]),
        "tsu_calibrate_copy": (ctypes.c_int, [vp, i64, P(ctypes.c_double)]),
        "tsu_sync": (ctypes.c_int, [vp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def strerror(status: int) -> str:
    return library().tsu_strerror(status).decode()


def _i32p(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))


# ---------------------------------------------------------------------------
# datasets (host side; no GPU needed)
# ---------------------------------------------------------------------------
class Dataset:
    """Pairs held by the library: read from sequences.txt, generated, or given."""

    def __init__(self, handle: ctypes.c_void_p):
        self._h = handle
        self._lib = library()

    @classmethod
    def read(cls, path: str) -> "Dataset":
        lib = library()
        h = ctypes.c_void_p()
        err = ctypes.create_string_buffer(512)
        r = lib.tsu_dataset_read(os.fsencode(path), ctypes.byref(h), err, 512)
        if r != TSU_OK:
            raise TsuError(r, f"tsu_dataset_read({path}): {err.value.decode(errors='replace')}")
        return cls(h)

    @classmethod
    def generate(cls, n: int, seed: int, len_min: int, len_max: Optional[int] = None,
                 err_min: float = 0.0, err_max: Optional[float] = None, edit_mode: int = 0) -> "Dataset":
        lib = library()
        cfg = GenCfg(seed, len_min, len_min if len_max is None else len_max, err_min,
                     err_min if err_max is None else err_max, edit_mode, 0)
        h = ctypes.c_void_p()
        r = lib.tsu_dataset_generate(ctypes.byref(cfg), n, ctypes.byref(h))
        if r != TSU_OK:
            raise TsuError(r, "tsu_dataset_generate")
        return cls(h)

    @classmethod
    def from_pairs(cls, pairs: Sequence[Tuple[str, str]]) -> "Dataset":
        lib = library()
        n = len(pairs)
        pats = [p.encode() for p, _ in pairs]
        txts = [t.encode() for _, t in pairs]
        parr = (ctypes.c_char_p * max(n, 1))(*pats)
        tarr = (ctypes.c_char_p * max(n, 1))(*txts)
        pl = np.array([len(p) for p in pats] or [0], dtype=np.int32)
        tl = np.array([len(t) for t in txts] or [0], dtype=np.int32)
        h = ctypes.c_void_p()
        r = lib.tsu_dataset_from_strings(parr, _i32p(pl), tarr, _i32p(tl), n, ctypes.byref(h))
        if r != TSU_OK:
            raise TsuError(r, "tsu_dataset_from_strings")
        return cls(h)

    def write(self, path: str, error_rate_field: int = 0) -> None:
        r = self._lib.tsu_dataset_write(self._h, os.fsencode(path), error_rate_field)
        if r != TSU_OK:
            raise TsuError(r, f"tsu_dataset_write({path})")

    def __len__(self) -> int:
        return int(self._lib.tsu_dataset_size(self._h))

    def pair(self, i: int) -> Tuple[str, str]:
        pp, tp = ctypes.c_void_p(), ctypes.c_void_p()
        pl, tl = ctypes.c_int32(), ctypes.c_int32()
        r = self._lib.tsu_dataset_pair(self._h, i, ctypes.byref(pp), ctypes.byref(pl), ctypes.byref(tp), ctypes.byref(tl))
        if r != TSU_OK:
            raise TsuError(r, "tsu_dataset_pair")
        return (ctypes.string_at(pp, pl.value).decode(), ctypes.string_at(tp, tl.value).decode())

    def pairs(self) -> List[Tuple[str, str]]:
        return [self.pair(i) for i in range(len(self))]

    def header(self) -> Tuple[int, int, int]:
        a, b, c = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
        self._lib.tsu_dataset_header(self._h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c))
        return a.value, b.value, c.value

    def lengths(self) -> Tuple[np.ndarray, np.ndarray]:
        n = len(self)
        pl = np.empty(max(n, 1), np.int32)
        tl = np.empty(max(n, 1), np.int32)
        r = self._lib.tsu_dataset_lengths(self._h, _i32p(pl), _i32p(tl))
        if r != TSU_OK:
            raise TsuError(r, "tsu_dataset_lengths")
        return pl[:n], tl[:n]

    def packed(self) -> Tuple[np.ndarray, np.ndarray]:
        """(words, pairs): every pair packed with tsu_pack into one arena,
        pattern then text, in order; pairs is a PAIR_DTYPE array."""
        n = len(self)
        pl, tl = self.lengths()
        nwp = (pl.astype(np.int64) + 15) // 16
        nwt = (tl.astype(np.int64) + 15) // 16
        pairs = np.zeros(n, PAIR_DTYPE)
        ends = np.cumsum(nwp + nwt)
        pairs["p_off"] = ends - (nwp + nwt)
        pairs["t_off"] = pairs["p_off"] + nwp
        pairs["p_len"], pairs["t_len"] = pl, tl
        words = np.zeros(int(ends[-1]) if n else 1, np.uint32)
        for i in range(n):
            p, t = self.pair(i)
            words[pairs["p_off"][i]:pairs["p_off"][i] + nwp[i]] = pack(p)[:nwp[i]]
            words[pairs["t_off"][i]:pairs["t_off"][i] + nwt[i]] = pack(t)[:nwt[i]]
        return words, pairs

    def close(self) -> None:
        if self._h:
            self._lib.tsu_dataset_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _out_scores(out: Optional[np.ndarray], n: int) -> np.ndarray:
    if out is None:
        return np.zeros(max(n, 1), np.int32)
    if out.dtype != np.int32 or out.ndim != 1 or not out.flags.c_contiguous or len(out) < max(n, 1):
        raise ValueError(f"out must be a contiguous int32 array of at least {max(n, 1)} entries")
    return out


def pack(seq: str) -> np.ndarray:
    lib = library()
    n = int(lib.tsu_packed_words(len(seq)))
    out = np.zeros(max(n, 1), dtype=np.uint32)
    r = lib.tsu_pack(seq.encode(), len(seq), out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)))
    if r != TSU_OK:
        raise TsuError(r, "tsu_pack")
    return out[:n]


def cigar_string(runs: Iterable[int]) -> str:
    return "".join(f"{r >> 2}{OPS[r & 3]}" for r in runs)


# ---------------------------------------------------------------------------
# GPU context
# ---------------------------------------------------------------------------
class Batch:
    def __init__(self, aligner: "Aligner", handle: ctypes.c_void_p):
        self._a = aligner
        self._h = handle

    def __len__(self) -> int:
        return int(library().tsu_batch_size(self._h))

    @property
    def nbytes(self) -> int:
        return int(library().tsu_batch_bytes(self._h))

    def close(self) -> None:
        if self._h:
            library().tsu_batch_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Aligner:
    """One device context with fixed gap-affine penalties (x, o, e)."""

    def __init__(self, mismatch: int = 4, gap_opening: int = 6, gap_extension: int = 2, device: int = 0,
                 gap_opening2: Optional[int] = None, gap_extension2: Optional[int] = None):
        """Gap-affine penalties; with gap_opening2 / gap_extension2 the 2-piece
        model, where a gap of length L costs min(o1 + L e1, o2 + L e2)."""
        self._lib = library()
        h = ctypes.c_void_p()
        if gap_opening2 is None and gap_extension2 is None:
            r = self._lib.tsu_create(ctypes.byref(h), device, mismatch, gap_opening, gap_extension)
            self.penalties = (mismatch, gap_opening, gap_extension)
        else:
            r = self._lib.tsu_create_2p(ctypes.byref(h), device, mismatch, gap_opening, gap_extension,
                                        gap_opening2, gap_extension2)
            self.penalties = (mismatch, gap_opening, gap_extension, gap_opening2, gap_extension2)
        if r != TSU_OK:
            raise TsuError(r, f"tsu_create: {strerror(r)}")
        self._h = h
        self._batches = weakref.WeakSet()

    def _check(self, r: int, what: str) -> None:
        if r != TSU_OK:
            raise TsuError(r, f"{what}: {self._lib.tsu_last_error(self._h).decode(errors='replace')}")

    def device_info(self) -> dict:
        cu, lds = ctypes.c_int32(), ctypes.c_int32()
        name = ctypes.create_string_buffer(256)
        self._check(self._lib.tsu_device_info(self._h, ctypes.byref(cu), ctypes.byref(lds), name, 256), "tsu_device_info")
        return {"cu_count": cu.value, "lds_bytes": lds.value, "name": name.value.decode()}

    def batch(self, data, first: int = 0, count: Optional[int] = None) -> Batch:
        ds = data if isinstance(data, Dataset) else Dataset.from_pairs(list(data))
        if count is None:
            count = len(ds) - first
        h = ctypes.c_void_p()
        self._check(self._lib.tsu_batch_from_dataset(self._h, ds._h, first, count, ctypes.byref(h)), "tsu_batch_from_dataset")
        b = Batch(self, h)
        self._batches.add(b)
        return b

    def align_plain(self, batch: Batch, wf_length: int, raise_on_error: bool = True, out: Optional[np.ndarray] = None):
        """Scores of every pair (plain form).  out: an int32 array of at least
        len(batch) entries to write them to (returned as its first len(batch))."""
        n = len(batch)
        scores = _out_scores(out, n)
        t = Timing()
        r = self._lib.tsu_align_plain(self._h, batch._h, wf_length, _i32p(scores), ctypes.byref(t))
        if raise_on_error:
            self._check(r, "tsu_align_plain")
        return scores[:n], t.as_dict()

    def align_fixed(self, patterns: Sequence[str], texts: Sequence[str], seq_len: int, wf_length: int):
        n = len(patterns)
        pbuf = b"".join(p.encode()[:seq_len].ljust(seq_len, b"A") for p in patterns)
        tbuf = b"".join(t.encode()[:seq_len].ljust(seq_len, b"A") for t in texts)
        scores = np.zeros(max(n, 1), np.int32)
        t = Timing()
        self._check(self._lib.tsu_align_fixed(self._h, pbuf, tbuf, n, seq_len, wf_length, _i32p(scores), ctypes.byref(t)),
                    "tsu_align_fixed")
        return scores[:n], t.as_dict()

    def align_batched(self, batch: Batch, out: Optional[np.ndarray] = None):
        """Scores of every pair (batched form); out as for align_plain."""
        n = len(batch)
        scores = _out_scores(out, n)
        t = Timing()
        self._check(self._lib.tsu_align_batched(self._h, batch._h, _i32p(scores), ctypes.byref(t)), "tsu_align_batched")
        return scores[:n], t.as_dict()

    def autotune(self, batch: Batch, sample_pairs: int = 4096) -> List[dict]:
        self._check(self._lib.tsu_autotune(self._h, batch._h, sample_pairs), "tsu_autotune")
        return self.tune_table()

    def tune_table(self) -> List[dict]:
        arr = (TuneEntry * 64)()
        n = ctypes.c_int32()
        self._check(self._lib.tsu_tune_table(self._h, arr, 64, ctypes.byref(n)), "tsu_tune_table")
        return [{"waves_per_pair": arr[i].waves_per_pair, "capacity": arr[i].capacity,
                 "bucket_bound": arr[i].bucket_bound} for i in range(min(n.value, 64))]

    def set_tune_table(self, entries: Sequence[dict]) -> None:
        """Replaces the per-bucket table (entries as tune_table returns
        them; an empty list clears it)."""
        arr = (TuneEntry * max(len(entries), 1))()
        for i, e in enumerate(entries):
            arr[i].waves_per_pair = int(e.get("waves_per_pair", 0))
            arr[i].capacity = int(e.get("capacity", 0))
            arr[i].bucket_bound = int(e["bucket_bound"])
        self._check(self._lib.tsu_set_tune_table(self._h, arr, len(entries)), "tsu_set_tune_table")

    CIGAR_AUTO, CIGAR_RECORDS, CIGAR_CHECKPOINTS = 0, 1, 2

    def set_cigar_memory(self, mode: int = 0, checkpoint_every: int = 0, budget_bytes: int = 0) -> None:
        """Backtrace memory of align_cigar: CIGAR_AUTO (records unless they
        do not fit one pair per CU within the budget), CIGAR_RECORDS or
        CIGAR_CHECKPOINTS (ring snapshots every checkpoint_every scores, 0 =
        chosen per launch).  budget_bytes: device memory one launch may hold
        for backtrace state, 0 = half the free memory."""
        self._check(self._lib.tsu_set_cigar_memory(self._h, mode, checkpoint_every, budget_bytes),
                    "tsu_set_cigar_memory")

    def set_poll_limit(self, polls: int = 0) -> None:
        """Polls of a neighbour's progress before a pair fails with
        TSU_EINTERNAL (score-only launches of 1,025-16,384 diagonals); 0 =
        the default."""
        self._check(self._lib.tsu_set_poll_limit(self._h, polls), "tsu_set_poll_limit")

    def align_cigar(self, batch: Batch, strings=True, count_runs: bool = False):
        """(scores, CIGAR strings, timing).  With strings=False the CIGARs are
        computed and copied to the host by the library as usual but not
        turned into Python strings (None is returned in their place); an
        integer m returns the strings of the first m pairs only.  The
        timing dict also holds the backtrace state: state_bytes (largest of
        one launch), record_launches, checkpoint_launches and kernels (the
        set of KERNEL_BITS names of the kernels that recorded); with
        count_runs, cigar_runs is the number of runs over all pairs (one
        library call per pair, so not for timed calls)."""
        n = len(batch)
        scores = np.zeros(max(n, 1), np.int32)
        t = Timing()
        h = ctypes.c_void_p()
        self._check(self._lib.tsu_align_cigar(self._h, batch._h, _i32p(scores), ctypes.byref(h), ctypes.byref(t)),
                    "tsu_align_cigar")
        count = n if strings is True else 0 if strings is False else min(int(strings), n)
        cigars = [] if count or strings is True else None
        try:
            rp = ctypes.POINTER(ctypes.c_uint32)()
            nr = ctypes.c_int32()
            for i in range(count):
                self._check(self._lib.tsu_cigars_get(h, i, ctypes.byref(rp), ctypes.byref(nr)), "tsu_cigars_get")
                cigars.append(cigar_string(rp[j] for j in range(nr.value)))
            runs = None
            if count_runs:
                runs = 0
                for i in range(n):
                    self._check(self._lib.tsu_cigars_get(h, i, ctypes.byref(rp), ctypes.byref(nr)), "tsu_cigars_get")
                    runs += nr.value
            sb, rl, cl = ctypes.c_int64(), ctypes.c_int32(), ctypes.c_int32()
            self._check(self._lib.tsu_cigars_stats(h, ctypes.byref(sb), ctypes.byref(rl), ctypes.byref(cl)),
                        "tsu_cigars_stats")
            km = ctypes.c_uint32()
            self._check(self._lib.tsu_cigars_kernels(h, ctypes.byref(km)), "tsu_cigars_kernels")
        finally:
            self._lib.tsu_cigars_free(h)
        td = t.as_dict()
        td.update(state_bytes=sb.value, record_launches=rl.value, checkpoint_launches=cl.value,
                  kernels={name for bit, name in KERNEL_BITS.items() if km.value & bit})
        if count_runs:
            td["cigar_runs"] = runs
        return scores[:n], cigars, td

    def calibrate_copy(self, nbytes: int) -> float:
        ms = ctypes.c_double()
        self._check(self._lib.tsu_calibrate_copy(self._h, nbytes, ctypes.byref(ms)), "tsu_calibrate_copy")
        return ms.value

    def sync(self) -> None:
        self._check(self._lib.tsu_sync(self._h), "tsu_sync")

    def close(self) -> None:
        if self._h:
            for b in list(self._batches):
                b.close()
            self._lib.tsu_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------------------
# serving
# ---------------------------------------------------------------------------
class Server:
    """Continuous-batching server on one Aligner's device (tsu_serve_*)."""

    def __init__(self, aligner: "Aligner", hbm_budget_bytes: int = 8 << 30, max_inflight: int = 4,
                 max_batch_pairs: int = 1024):
        self._lib = library()
        self._a = aligner
        cfg = ServeCfg(hbm_budget_bytes, max_inflight, max_batch_pairs)
        h = ctypes.c_void_p()
        aligner._check(self._lib.tsu_serve_open(aligner._h, ctypes.byref(cfg), ctypes.byref(h)), "tsu_serve_open")
        self._h = h

    def submit(self, pairs: Sequence[Tuple[str, str]]) -> int:
        n = len(pairs)
        pats = [p.encode() for p, _ in pairs]
        txts = [t.encode() for _, t in pairs]
        parr = (ctypes.c_char_p * max(n, 1))(*pats)
        tarr = (ctypes.c_char_p * max(n, 1))(*txts)
        pl = np.array([len(p) for p in pats] or [0], dtype=np.int32)
        tl = np.array([len(t) for t in txts] or [0], dtype=np.int32)
        ticket = ctypes.c_uint64()
        r = self._lib.tsu_serve_submit(self._h, parr, _i32p(pl), tarr, _i32p(tl), n, ctypes.byref(ticket))
        if r != TSU_OK:
            raise TsuError(r, f"tsu_serve_submit: {strerror(r)}")
        return ticket.value

    def submit_dataset(self, ds: "Dataset", first: int, count: int) -> int:
        """Submits pairs [first, first+count) of ds without copying them into Python."""
        lib = self._lib
        parr = (ctypes.c_char_p * max(count, 1))()
        tarr = (ctypes.c_char_p * max(count, 1))()
        pl = np.zeros(max(count, 1), np.int32)
        tl = np.zeros(max(count, 1), np.int32)
        pp, tp = ctypes.c_void_p(), ctypes.c_void_p()
        a, b = ctypes.c_int32(), ctypes.c_int32()
        for j in range(count):
            lib.tsu_dataset_pair(ds._h, first + j, ctypes.byref(pp), ctypes.byref(a), ctypes.byref(tp), ctypes.byref(b))
            ctypes.cast(parr, ctypes.POINTER(ctypes.c_void_p))[j] = pp.value
            ctypes.cast(tarr, ctypes.POINTER(ctypes.c_void_p))[j] = tp.value
            pl[j], tl[j] = a.value, b.value
        ticket = ctypes.c_uint64()
        r = lib.tsu_serve_submit(self._h, parr, _i32p(pl), tarr, _i32p(tl), count, ctypes.byref(ticket))
        if r != TSU_OK:
            raise TsuError(r, f"tsu_serve_submit: {strerror(r)}")
        return ticket.value

    def submit_packed(self, words: np.ndarray, pairs: np.ndarray) -> int:
        """Submits pairs already packed: `pairs` is a structured array with the
        fields of tsu_pair (p_off, t_off, p_len, t_len) indexing `words`."""
        words = np.ascontiguousarray(words, dtype=np.uint32)
        pairs = np.ascontiguousarray(pairs, dtype=PAIR_DTYPE)
        ticket = ctypes.c_uint64()
        r = self._lib.tsu_serve_submit_packed(self._h, words.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)),
                                              len(words), pairs.ctypes.data_as(ctypes.POINTER(Pair)), len(pairs),
                                              ctypes.byref(ticket))
        if r != TSU_OK:
            raise TsuError(r, f"tsu_serve_submit_packed: {strerror(r)}")
        return ticket.value

    def _result(self, fn, ticket: int, n: int):
        scores = np.zeros(max(n, 1), np.int32)
        st = ReqStats()
        r = fn(self._h, ticket, _i32p(scores), ctypes.byref(st))
        return r, scores[:st.n_pairs if r != TSU_EAGAIN else 0], {f: getattr(st, f) for f, _ in ReqStats._fields_}

    def poll(self, ticket: int, n: int):
        """(status, scores, stats); status TSU_EAGAIN while not done."""
        return self._result(self._lib.tsu_serve_poll, ticket, n)

    def wait(self, ticket: int, n: int):
        return self._result(self._lib.tsu_serve_wait, ticket, n)

    def close(self) -> int:
        r = TSU_OK
        if self._h:
            r = self._lib.tsu_serve_close(self._h)
            self._h = None
        return r

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
