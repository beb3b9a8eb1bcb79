"""The C-ABI library builds for gfx950, loads, and exports every function include/diffractom_hip.h declares; argument
errors come back as status codes with a message (no GPU needed: the checks run before any HIP call)."""
import ctypes
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]


def _declared():
    text = (ROOT / "include" / "diffractom_hip.h").read_text()
    return sorted(set(re.findall(r"\b(dfm_[a-z0-9_]+)\s*\(", text)))


@pytest.fixture(scope="module")
def lib():
    from diffractom_amd import _lib
    if not _lib.LIB_PATH.exists():
        from diffractom_amd import build
        build.build(verbose=False)
    return _lib.load()


def test_every_declared_symbol_is_exported(lib):
    names = _declared()
    assert len(names) > 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_bindings_cover_the_header():
    from diffractom_amd import _lib
    names = set(_declared()) - {"dfm_version", "dfm_last_error"}
    assert names <= set(_lib.SIGNATURES), sorted(names - set(_lib.SIGNATURES))


def test_argument_errors_are_status_codes(lib):
    from diffractom_amd import _lib
    assert lib.dfm_version() >= 100
    with pytest.raises(ValueError, match="null pointer"):
        _lib.call("dfm_gather_k", None, None, 10, 0, 4, 4, None)
    with pytest.raises(ValueError, match="multiples of 4"):
        _lib.call("dfm_radon_forward", ctypes.c_void_p(8), ctypes.c_void_p(8), ctypes.c_void_p(8), 4, 4, 5, 3, 6, 0,
                  6, 6, 1.0, 2, None)
    with pytest.raises(ValueError, match="ntap"):
        _lib.call("dfm_radon_backward", ctypes.c_void_p(8), ctypes.c_void_p(8), ctypes.c_void_p(8), 4, 4, 5, 3, 8,
                  0, 8, 8, 7, None)
