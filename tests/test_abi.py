"""The C-ABI library loads and exports every entry point include/tnhip.h declares
(CPU-only: no compute call is made)."""
import ctypes
import re

from conftest import PKG, ROOT


def declared_symbols():
    text = (ROOT / "include" / "tnhip.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b(tnh_[a-z0-9_]+)\s*\(", text)
    return sorted(set(names))


def test_header_declares_entry_points():
    names = declared_symbols()
    for must in ["tnh_ctx_create", "tnh_heff_apply", "tnh_env_left", "tnh_env_right",
                 "tnh_eigsolve_ground", "tnh_svd_trunc", "tnh_replacesites2", "tnh_gemm"]:
        assert must in names


def test_library_exports_all_declared_symbols():
    lib = ctypes.CDLL(str(PKG / "tnhip" / "libtnhip.so"))
    missing = [n for n in declared_symbols() if not hasattr(lib, n)]
    assert not missing, f"declared but not exported: {missing}"


def test_python_binding_covers_header():
    import tnhip._lib as L

    assert not L.MISSING, f"binding lists symbols the library lacks: {L.MISSING}"
    unbound = [n for n in declared_symbols() if n not in L.SIGNATURES]
    assert not unbound, f"declared in the header but not bound in Python: {unbound}"


def test_version_string():
    import tnhip

    assert "gfx950" in tnhip.version()
