"""The C-ABI library builds/loads and exports every entry point include/*.h declares
(no device calls: runs without a GPU)."""
import os
import re

from dmtz_amd import _native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    names = set()
    for h in ("dmtz.h", "dmtz_learned.h"):
        src = open(os.path.join(ROOT, "include", h)).read()
        names |= set(re.findall(r"\b(dmtz_\w+)\s*\(", src))
    return sorted(names)


def test_header_symbols_exported():
    L = _native.lib()
    names = declared()
    assert len(names) >= 15
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert set(names) == set(_native.EXPORTS)
