"""The C ABI library loads and exports every entry point its headers
declare (no device work: runs without a GPU)."""
import ctypes
import glob
import os
import re

from volcanor_amd import vlm

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names |= set(re.findall(r"\b(vlm_[a-z0-9_]+)\s*\(", text))
    return names


def test_header_symbols_exported():
    lib = ctypes.CDLL(vlm.LIB_PATH)
    names = declared_symbols()
    assert len(names) > 30
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_header():
    names = declared_symbols()
    bound = set(vlm._SIGS) | {"vlm_last_error"}
    assert names <= bound, sorted(names - bound)


def test_null_context_is_an_error_not_a_crash():
    lib = vlm.load_library()
    assert lib.vlm_set_precision(None, 0) != 0
    assert lib.vlm_destroy(None) == 0
    assert lib.vlm_last_error(None) == b"null context"
