"""The C-ABI library loads and exports every entry point include/codon_nsga3.h
declares (no GPU needed: no compute call is made)."""
import ctypes
import os
import re
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    text = open(os.path.join(ROOT, "include", "codon_nsga3.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|int64_t|const char \*)\s*\*?\s*(cd_\w+)\s*\(", text, re.M)))


def test_header_matches_binding(cdp):
    assert declared() == sorted(cdp.SYMBOLS)


def test_library_exports_every_symbol(cdp):
    assert os.path.exists(cdp.LIB_PATH), "build first: python -c 'import __graft_entry__ as g; g.build()'"
    out = subprocess.run(["nm", "-D", "--defined-only", cdp.LIB_PATH], capture_output=True, text=True, check=True)
    exported = set(re.findall(r" T (cd_\w+)$", out.stdout, re.M))
    assert set(declared()) <= exported
    lib = ctypes.CDLL(cdp.LIB_PATH)
    for name in declared():
        assert getattr(lib, name) is not None


def test_device_bytes_is_host_only(cdp):
    # pure size arithmetic, callable without a device
    b1 = cdp.device_bytes(501, 4, 1024, 5)
    b2 = cdp.device_bytes(1501, 8, 4096, 5)
    assert 0 < b1 < b2
    assert b2 > 2 * 8192 * 8 * 1501  # two genome pools at least
    assert cdp.device_bytes(2, 4, 1024, 5) == -1
