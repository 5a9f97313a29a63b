"""ctypes wrapper of oracle/liboracle.so -- TEST INFRASTRUCTURE ONLY.

Imported by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg,
always to check results (or as the timed CPU baseline), never by the product.
See wfa_oracle.c for the algorithm and the reference lines it follows.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from typing import List, Optional, Sequence, Tuple

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
REF_GENERATOR = os.path.join(HERE, "_ref", "generate_dataset")


class Counters(ctypes.Structure):
    _fields_ = [("cells", ctypes.c_int64), ("extend_words", ctypes.c_int64),
                ("steps", ctypes.c_int64), ("max_width", ctypes.c_int64),
                ("min_lo", ctypes.c_int64), ("max_hi", ctypes.c_int64)]

    def as_dict(self) -> dict:
        return {f: int(getattr(self, f)) for f, _ in self._fields_}


_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE, "liboracle.so"], check=True)


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        l = ctypes.CDLL(LIB)
        l.orc_gotoh.restype = ctypes.c_int64
        l.orc_gotoh.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_char_p, ctypes.c_int,
                                ctypes.c_int, ctypes.c_int, ctypes.c_int]
        l.orc_wfa.restype = ctypes.c_int
        l.orc_wfa.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_char_p, ctypes.c_int,
                              ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                              ctypes.c_char_p, ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                              ctypes.POINTER(Counters)]
        P = ctypes.POINTER
        l.orc_wfa_nulled.restype = ctypes.c_int
        l.orc_wfa_nulled.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_char_p, ctypes.c_int,
                                     ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                     ctypes.c_int, ctypes.c_char_p, ctypes.c_int, P(ctypes.c_int)]
        l.orc_wfa_batch.restype = ctypes.c_int
        l.orc_wfa_batch.argtypes = [ctypes.c_char_p, P(ctypes.c_int64), P(ctypes.c_int32), P(ctypes.c_int64),
                                    P(ctypes.c_int32), ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                    ctypes.c_int, P(ctypes.c_int32), ctypes.c_int, ctypes.c_char_p,
                                    P(ctypes.c_int64), P(ctypes.c_int32), P(Counters)]
        l.orc_max_threads.restype = ctypes.c_int
        l.orc_wfa_prio.restype = ctypes.c_int
        l.orc_wfa_prio.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_char_p, ctypes.c_int,
                                   ctypes.c_int, ctypes.c_int, ctypes.c_int, P(ctypes.c_int),
                                   ctypes.c_char_p, ctypes.c_int, P(ctypes.c_int)]
        l.orc_gotoh2p.restype = ctypes.c_int64
        l.orc_gotoh2p.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_char_p, ctypes.c_int] + [ctypes.c_int] * 5
        l.orc_wfa2p.restype = ctypes.c_int
        l.orc_wfa2p.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_char_p, ctypes.c_int] + [ctypes.c_int] * 5 + \
            [P(ctypes.c_int64)]
        l.orc_wfa2p_prio.restype = ctypes.c_int
        l.orc_wfa2p_prio.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_char_p, ctypes.c_int] + \
            [ctypes.c_int] * 5 + [P(ctypes.c_int), ctypes.c_char_p, ctypes.c_int, P(ctypes.c_int)]
        l.orc_wfa2p_batch.restype = ctypes.c_int
        l.orc_wfa2p_batch.argtypes = [ctypes.c_char_p, P(ctypes.c_int64), P(ctypes.c_int32), P(ctypes.c_int64),
                                      P(ctypes.c_int32), ctypes.c_int] + [ctypes.c_int] * 6 + \
            [P(ctypes.c_int32), P(ctypes.c_int64)]
        _lib = l
    return _lib


def gotoh(p: str, t: str, x: int = 4, o: int = 6, e: int = 2) -> int:
    return int(lib().orc_gotoh(p.encode(), len(p), t.encode(), len(t), x, o, e))


def wfa_nulled(p: str, t: str, x: int = 4, o: int = 6, e: int = 2, cigar: bool = False):
    """Score (and, with cigar=True, (score, ops)) with every out-of-bounds
    cell NULLed where it is computed and no trim (the neighbour-synchronised
    GPU kernel's semantics)."""
    cap = len(p) + len(t) + 1
    buf = ctypes.create_string_buffer(cap) if cigar else None
    n = ctypes.c_int(0)
    r = lib().orc_wfa_nulled(p.encode(), len(p), t.encode(), len(t), x, o, e, 1 if cigar else 0,
                             buf, cap if cigar else 0, ctypes.byref(n))
    if r < 0:
        raise RuntimeError(f"orc_wfa_nulled failed: {r}")
    return (r, buf.raw[:n.value].decode()) if cigar else r


def wfa(p: str, t: str, x: int = 4, o: int = 6, e: int = 2, cigar: bool = False) -> Tuple[int, Optional[str], dict]:
    """Score (and the operation string 'MXID...' when cigar) of one pair."""
    cap = len(p) + len(t) + 1
    buf = ctypes.create_string_buffer(cap) if cigar else None
    n = ctypes.c_int(0)
    c = Counters()
    r = lib().orc_wfa(p.encode(), len(p), t.encode(), len(t), x, o, e, 1 if cigar else 0,
                      buf, cap, ctypes.byref(n), ctypes.byref(c))
    if r < 0:
        raise RuntimeError(f"orc_wfa failed: {r}")
    ops = buf.raw[:n.value].decode() if cigar else None
    return r, ops, c.as_dict()


# Backtrace priorities of WFA2-lib (common.h:67-77), in the order the oracle
# takes them: M (mismatch), D_ext, D_open, I_ext, I_open.
PRIORITY = (9, 6, 5, 2, 1)
PRIORITY_KINDS = ("M", "D_ext", "D_open", "I_ext", "I_open")


def wfa_prio(p: str, t: str, x: int, o: int, e: int, prio: Sequence[int]) -> Tuple[int, str]:
    """(score, ops) with the backtrace priorities `prio` (order of PRIORITY_KINDS)."""
    cap = len(p) + len(t) + 1
    buf = ctypes.create_string_buffer(cap)
    n = ctypes.c_int(0)
    pr = (ctypes.c_int * 5)(*prio)
    r = lib().orc_wfa_prio(p.encode(), len(p), t.encode(), len(t), x, o, e, pr, buf, cap, ctypes.byref(n))
    if r < 0:
        raise RuntimeError(f"orc_wfa_prio failed: {r}")
    return r, buf.raw[:n.value].decode()


def gotoh2p(p: str, t: str, x: int, o1: int, e1: int, o2: int, e2: int) -> int:
    """2-piece Gotoh DP score (gap of length L costs min(o1 + L e1, o2 + L e2))."""
    return int(lib().orc_gotoh2p(p.encode(), len(p), t.encode(), len(t), x, o1, e1, o2, e2))


def wfa2p(p: str, t: str, x: int, o1: int, e1: int, o2: int, e2: int) -> Tuple[int, int]:
    """Gap-affine 2-piece WFA score and the number of computed cells."""
    c = ctypes.c_int64(0)
    r = lib().orc_wfa2p(p.encode(), len(p), t.encode(), len(t), x, o1, e1, o2, e2, ctypes.byref(c))
    if r < 0:
        raise RuntimeError(f"orc_wfa2p failed: {r}")
    return r, c.value


# backtrace type values of the 2-piece walk (common.h:67-77), in the order
# orc_wfa2p_prio takes them
PRIORITY_2P = (9, 8, 7, 6, 5, 4, 3, 2, 1)
PRIORITY_2P_KINDS = ("M", "D2_ext", "D2_open", "D1_ext", "D1_open", "I2_ext", "I2_open", "I1_ext", "I1_open")


def wfa2p_cigar(p: str, t: str, x: int, o1: int, e1: int, o2: int, e2: int,
                prio: Optional[Sequence[int]] = None) -> Tuple[int, str]:
    """Score and CIGAR ops (forward 'M'/'X'/'I'/'D' string) of the 2-piece WFA
    backtrace; prio overrides the nine type values (order of PRIORITY_2P_KINDS)."""
    cap = 2 * (len(p) + len(t)) + 16
    buf = ctypes.create_string_buffer(cap)
    n = ctypes.c_int(0)
    pr = None
    if prio is not None:
        pr = (ctypes.c_int * 9)(*prio)
    r = lib().orc_wfa2p_prio(p.encode(), len(p), t.encode(), len(t), x, o1, e1, o2, e2, pr, buf, cap, ctypes.byref(n))
    if r < 0:
        raise RuntimeError(f"orc_wfa2p_prio failed: {r}")
    return r, buf.raw[:n.value].decode()


def rescore2p(ops: str, p: str, t: str, x: int, o1: int, e1: int, o2: int, e2: int) -> int:
    """2-piece cost of an alignment given as forward ops (a gap run of length L
    costs min(o1 + L e1, o2 + L e2)); raises if the ops do not align p with t."""
    i = j = 0
    cost = 0
    for c in ops:
        if c == "M" or c == "X":
            if i >= len(p) or j >= len(t) or (p[i] == t[j]) != (c == "M"):
                raise ValueError(f"{c} at ({i}, {j}) does not fit the sequences")
            cost += x if c == "X" else 0
            i += 1; j += 1
        elif c == "I":
            j += 1
        elif c == "D":
            i += 1
        else:
            raise ValueError(f"unknown op {c!r}")
    if i != len(p) or j != len(t):
        raise ValueError("ops do not cover both sequences")
    pos = 0
    while pos < len(ops):
        end = pos
        while end < len(ops) and ops[end] == ops[pos]:
            end += 1
        if ops[pos] in "ID":
            n = end - pos
            cost += min(o1 + e1 * n, o2 + e2 * n)
        pos = end
    return cost


def wfa2p_batch(pairs: Sequence[Tuple[str, str]], x: int, o1: int, e1: int, o2: int, e2: int,
                threads: int = 0) -> Tuple[np.ndarray, int]:
    """Scores (np.int32) of many pairs under 2-piece penalties, and total cells."""
    n = len(pairs)
    chunks, p_off, p_len, t_off, t_len = [], [], [], [], []
    pos = 0
    for p, t in pairs:
        pb, tb = p.encode(), t.encode()
        p_off.append(pos); p_len.append(len(pb)); chunks.append(pb); pos += len(pb)
        t_off.append(pos); t_len.append(len(tb)); chunks.append(tb); pos += len(tb)
    blob = b"".join(chunks) or b"\0"
    arr = lambda v, dt: np.ascontiguousarray(np.array(v or [0], dtype=dt))
    po, pl, to, tl = arr(p_off, np.int64), arr(p_len, np.int32), arr(t_off, np.int64), arr(t_len, np.int32)
    scores = np.zeros(max(n, 1), np.int32)
    cells = ctypes.c_int64(0)
    P = ctypes.POINTER
    failed = lib().orc_wfa2p_batch(blob, po.ctypes.data_as(P(ctypes.c_int64)), pl.ctypes.data_as(P(ctypes.c_int32)),
                                   to.ctypes.data_as(P(ctypes.c_int64)), tl.ctypes.data_as(P(ctypes.c_int32)), n,
                                   x, o1, e1, o2, e2, threads, scores.ctypes.data_as(P(ctypes.c_int32)),
                                   ctypes.byref(cells))
    if failed:
        raise RuntimeError(f"orc_wfa2p_batch: {failed} pairs failed")
    return scores[:n], cells.value


def rle(ops: str) -> str:
    out = []
    i = 0
    while i < len(ops):
        j = i
        while j < len(ops) and ops[j] == ops[i]:
            j += 1
        out.append(f"{j - i}{ops[i]}")
        i = j
    return "".join(out)


def wfa_batch(pairs: Sequence[Tuple[str, str]], x: int = 4, o: int = 6, e: int = 2, threads: int = 0,
              cigar: bool = False):
    """Scores (np.int32), CIGAR op strings (or None) and summed counters."""
    n = len(pairs)
    chunks, p_off, p_len, t_off, t_len = [], [], [], [], []
    pos = 0
    for p, t in pairs:
        pb, tb = p.encode(), t.encode()
        p_off.append(pos); p_len.append(len(pb)); chunks.append(pb); pos += len(pb)
        t_off.append(pos); t_len.append(len(tb)); chunks.append(tb); pos += len(tb)
    blob = b"".join(chunks) or b"\0"
    arr = lambda v, dt: np.ascontiguousarray(np.array(v or [0], dtype=dt))
    po, pl, to, tl = arr(p_off, np.int64), arr(p_len, np.int32), arr(t_off, np.int64), arr(t_len, np.int32)
    scores = np.zeros(max(n, 1), np.int32)
    c = Counters()
    ops_buf = ops_off = ops_len = None
    if cigar:
        caps = (pl[:n].astype(np.int64) + tl[:n].astype(np.int64))
        ops_off = np.zeros(max(n, 1), np.int64)
        if n:
            ops_off[1:n] = np.cumsum(caps)[:-1]
        ops_buf = ctypes.create_string_buffer(int(caps.sum()) + 1)
        ops_len = np.zeros(max(n, 1), np.int32)
    P = ctypes.POINTER
    failed = lib().orc_wfa_batch(blob, po.ctypes.data_as(P(ctypes.c_int64)), pl.ctypes.data_as(P(ctypes.c_int32)),
                                 to.ctypes.data_as(P(ctypes.c_int64)), tl.ctypes.data_as(P(ctypes.c_int32)), n,
                                 x, o, e, threads, scores.ctypes.data_as(P(ctypes.c_int32)),
                                 1 if cigar else 0, ops_buf,
                                 ops_off.ctypes.data_as(P(ctypes.c_int64)) if cigar else None,
                                 ops_len.ctypes.data_as(P(ctypes.c_int32)) if cigar else None,
                                 ctypes.byref(c))
    if failed:
        raise RuntimeError(f"orc_wfa_batch: {failed} pairs failed")
    ops: Optional[List[str]] = None
    if cigar:
        raw = ops_buf.raw
        ops = [raw[ops_off[i]:ops_off[i] + ops_len[i]].decode() for i in range(n)]
    return scores[:n], ops, c.as_dict()


def max_threads() -> int:
    return int(lib().orc_max_threads())


def rescore(ops: str, p: str, t: str, x: int, o: int, e: int) -> int:
    """Cost of an operation string; raises if it does not align p with t."""
    i = j = 0
    cost = 0
    prev = None
    for c in ops:
        if c == "M":
            if p[i] != t[j]:
                raise ValueError(f"M at ({i},{j}) is not a match")
            i += 1; j += 1
        elif c == "X":
            if p[i] == t[j]:
                raise ValueError(f"X at ({i},{j}) is a match")
            i += 1; j += 1; cost += x
        elif c == "I":
            cost += e if prev == "I" else o + e
            j += 1
        elif c == "D":
            cost += e if prev == "D" else o + e
            i += 1
        else:
            raise ValueError(f"bad op {c!r}")
        prev = c
    if i != len(p) or j != len(t):
        raise ValueError("operations do not consume both sequences")
    return cost


def expand(cigar: str) -> str:
    """'3M1X' -> 'MMMX'."""
    out = []
    num = ""
    for ch in cigar:
        if ch.isdigit():
            num += ch
        else:
            out.append(ch * int(num))
            num = ""
    return "".join(out)
