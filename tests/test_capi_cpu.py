"""The C-ABI library loads without a GPU and exports every symbol the public
header declares; the Python binding declares argtypes for all of them."""
import ctypes
import re
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent


def declared_symbols():
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        text = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        names |= set(re.findall(r"\b(stp_[a-z0-9_]+)\s*\(", text))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    from seismic_engine import _lib
    lib = ctypes.CDLL(str(_lib.LIB_PATH))
    names = declared_symbols()
    assert len(names) > 10
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_header():
    from seismic_engine import _lib
    bound = set(_lib.SIGNATURES) | {"stp_last_error"}
    missing = [n for n in declared_symbols() if n not in bound]
    assert not missing, missing


def test_no_gpu_call_needed_to_load():
    from seismic_engine import _lib
    lib = _lib.lib()
    assert lib.stp_version() >= 100
    assert lib.stp_last_error() == b""
