"""The C-ABI library loads without a GPU and exports every function its public header declares."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER_DIR = os.path.join(ROOT, "include")
LIB = os.path.join(ROOT, "hero_amd", "_lib", "libhero_gpu.so")

_DECL = re.compile(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b(hero_\w+)\s*\(", re.M)


def declared_functions():
    names = []
    for fn in sorted(os.listdir(HEADER_DIR)):
        if not fn.endswith(".h"):
            continue
        text = open(os.path.join(HEADER_DIR, fn)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        text = re.sub(r"//[^\n]*", "", text)
        names += [(fn, m.group(1)) for m in _DECL.finditer(text)]
    return names


def test_header_declares_functions():
    names = [n for _, n in declared_functions()]
    assert "hero_create" in names and "hero_band_align_batch" in names
    assert len(names) == len(set(names)), "a function is declared twice"


@pytest.mark.skipif(not os.path.exists(LIB), reason="libhero_gpu.so is not built")
def test_library_exports_every_declared_function():
    lib = ctypes.CDLL(LIB)
    missing = [f"{h}:{n}" for h, n in declared_functions() if not hasattr(lib, n)]
    assert not missing, f"declared but not exported: {missing}"


@pytest.mark.skipif(not os.path.exists(LIB), reason="libhero_gpu.so is not built")
def test_default_options_and_null_context_errors():
    """Host-only entry points: defaults follow the reference's, and a null context is an argument error."""
    from hero_amd import gpu

    o = gpu.default_options()
    assert (o.k, o.hash_factor, o.min_overlap) == (39, 32, 40)  # Config.cpp:17,32; HashTableLongKey.cpp:41-68
    assert (o.hit, o.miss, o.dx, o.dy, o.ex, o.sx) == (1, -4, -4, -4, -1, 0)  # BandAligner.cpp:528-542
    assert abs(o.consensus_freq - 0.6) < 1e-12
    lib = ctypes.CDLL(LIB)
    lib.hero_last_error.restype = ctypes.c_char_p
    assert lib.hero_build_index(None) == -1


def test_table_size_list_matches_reference():
    """hero_prime_larger_than against getPrimeLargerThanNumber (src/HashTableLongKey.cpp:101-166) at
    every list entry and its neighbours, random sizes and sizes beyond the list (n + 1)."""
    from helpers import load_jsonl
    from hero_amd import gpu
    f = gpu.lib().hero_prime_larger_than
    bad = [(c["n"], c["P"], f(c["n"])) for c in load_jsonl("prime_cases.jsonl") if f(c["n"]) != c["P"]]
    assert not bad, bad[:5]
