"""CPU-side checks of the C ABI library: it loads and exports every symbol the
public header declares (no compute calls: there is no GPU here)."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_declared_symbol():
    from odemi import _lib
    L = ctypes.CDLL(_lib.LIB_PATH)
    hdr = open(os.path.join(ROOT, "include", "odemi.h")).read()
    hdr = re.sub(r"/\*.*?\*/", "", hdr, flags=re.S)
    names = set(re.findall(r"\b(odemi_[a-z_0-9]+)\s*\(", hdr))
    assert names, "no declarations found"
    missing = [n for n in sorted(names) if not hasattr(L, n)]
    assert not missing, missing
