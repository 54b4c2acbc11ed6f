"""CPU tests: the C-ABI library loads and exports every function include/sgfim.h declares."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "sgfim.h")
LIB = os.path.join(ROOT, "20261008-self-growing-evidence-agent_amd", "sgfim", "libsgfim.so")


def _declared():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(sgfim_[a-z_0-9]+)\s*\(", txt)))


def test_header_declares_entry_points():
    names = _declared()
    for n in ("sgfim_create", "sgfim_load_csr", "sgfim_count_batch", "sgfim_pair_matrix", "sgfim_mine",
              "sgfim_anytime_begin", "sgfim_anytime_step", "sgfim_result_get"):
        assert n in names


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(LIB)
    missing = [n for n in _declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_version_string():
    lib = ctypes.CDLL(LIB)
    lib.sgfim_version.restype = ctypes.c_char_p
    assert b"gfx950" in lib.sgfim_version()
