"""The C-ABI libraries load and export every function their headers declare
(no GPU work is done here)."""
import ctypes
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent


def declared(header: str) -> list[str]:
    text = (ROOT / "include" / header).read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b((?:fhip|fhost)_[a-z0-9_]+)\s*\(", text)))


@pytest.mark.parametrize("header,lib", [("flurry_hip.h", "libflurry_hip.so"), ("flurry_host.h", "libflurry_host.so")])
def test_exports(header, lib):
    names = declared(header)
    assert len(names) > 10
    so = ctypes.CDLL(str(ROOT / "flurrypp_amd" / "lib" / lib))
    missing = [n for n in names if not hasattr(so, n)]
    assert not missing, f"{lib} lacks {missing}"


def test_hip_library_has_gfx950_code():
    data = (ROOT / "flurrypp_amd" / "lib" / "libflurry_hip.so").read_bytes()
    assert b"gfx950" in data


def test_errors_without_context():
    """Entry points reject a null context with FHIP_EINVAL instead of crashing."""
    so = ctypes.CDLL(str(ROOT / "flurrypp_amd" / "lib" / "libflurry_hip.so"))
    so.fhip_last_error.restype = ctypes.c_char_p
    assert so.fhip_calc_residual(None, 0) == 1
    assert b"null" in so.fhip_last_error()
    assert so.fhip_update(None, 1, 0) == 1
    assert so.fhip_create(None, None) == 1
