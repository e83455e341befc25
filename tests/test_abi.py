"""The C-ABI library loads and exports every entry point include/bzcull.h declares (no GPU needed)."""
import re
import subprocess
from pathlib import Path

from blitzen_amd import _capi

ROOT = Path(__file__).resolve().parents[1]


def declared_symbols():
    text = (ROOT / "include" / "bzcull.h").read_text()
    return sorted(set(re.findall(r"^BZ_API\s+[\w\s\*]+?\b(bz_\w+)\s*\(", text, flags=re.M)))


def test_header_declarations_found():
    syms = declared_symbols()
    assert len(syms) >= 50
    assert "bz_cull" in syms and "bz_pyramid_build" in syms and "bz_scene_create" in syms


def test_library_exports_every_declared_symbol():
    lib = _capi.load()
    out = subprocess.run(["nm", "-D", "--defined-only", str(_capi.LIB_PATH)], capture_output=True, text=True).stdout
    exported = set(re.findall(r"\bT (bz_\w+)", out))
    missing = [s for s in declared_symbols() if s not in exported]
    assert not missing, missing
    for s in declared_symbols():
        assert hasattr(lib, s)


def test_binding_covers_every_declared_symbol():
    assert sorted(_capi.SIGNATURES) == declared_symbols()


def test_no_device_calls_without_gpu():
    lib = _capi.load()
    assert lib.bz_abi_version() == 2
    assert lib.bz_status_str(0) == b"ok"
    assert lib.bz_status_str(-4) == b"output capacity exceeded"
