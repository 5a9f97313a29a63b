"""The C-ABI library loads and exports every symbol include/tsunami_wfa.h declares.

No compute call is made here (no GPU needed)."""
import ctypes
import os
import re
import subprocess


def header_symbols(repo):
    with open(os.path.join(repo, "include", "tsunami_wfa.h")) as f:
        src = f.read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(tsu_[a-z0-9_]+)\s*\(", src)))


def test_header_matches_binding_list(repo, tsu):
    syms = header_symbols(repo)
    assert len(syms) >= 30
    assert sorted(tsu.EXPORTS) == syms


def test_library_exports_every_symbol(repo, tsu):
    lib = ctypes.CDLL(tsu.LIB_PATH)
    for s in header_symbols(repo):
        assert hasattr(lib, s), s
    out = subprocess.run(["nm", "-D", "--defined-only", tsu.LIB_PATH], capture_output=True, text=True, check=True).stdout
    defined = set(l.split()[-1] for l in out.splitlines() if l.strip())
    missing = [s for s in header_symbols(repo) if s not in defined]
    assert not missing, missing


def test_library_targets_gfx950(tsu):
    data = open(tsu.LIB_PATH, "rb").read()
    assert b"gfx950" in data


def test_host_only_entry_points(tsu):
    lib = tsu.library()
    assert lib.tsu_packed_words(0) == 0
    assert lib.tsu_packed_words(16) == 1
    assert lib.tsu_packed_words(17) == 2
    assert lib.tsu_strerror(tsu.TSU_EKOVERFLOW).decode() == "diagonal range exceeded wf_length"
    # invalid penalties are rejected before any device call
    h = ctypes.c_void_p()
    assert lib.tsu_create(ctypes.byref(h), 0, 0, 6, 2) == tsu.TSU_EINVAL
    assert lib.tsu_create(ctypes.byref(h), 0, 4, -1, 2) == tsu.TSU_EINVAL
    assert lib.tsu_create(ctypes.byref(h), 0, 4, 6, 0) == tsu.TSU_EINVAL
