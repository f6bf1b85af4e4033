"""The C-ABI library loads, exports every symbol include/slemm_gpu.h declares, and reports
argument errors through status codes without touching a GPU."""
import ctypes as C
import re
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
HEADER = ROOT / "include" / "slemm_gpu.h"


def declared_symbols():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(slemm_gpu_\w+)\s*\(", text)))


@pytest.fixture(scope="module")
def lib():
    from slemm_amd import _lib
    return _lib.load()


def test_header_declares_the_abi():
    syms = declared_symbols()
    assert "slemm_gpu_op" in syms and "slemm_gpu_lseed" in syms and "slemm_gpu_sldf_reml" in syms
    assert len(syms) >= 25


def test_library_exports_every_declared_symbol(lib):
    from slemm_amd import _lib
    out = subprocess.run(["nm", "-D", "--defined-only", str(_lib.LIB_PATH)], capture_output=True, text=True,
                         check=True).stdout
    exported = set(re.findall(r"\bT (slemm_gpu_\w+)", out))
    missing = [s for s in declared_symbols() if s not in exported]
    assert not missing, missing
    assert set(declared_symbols()) == set(_lib.SIGNATURES), "ctypes signatures out of sync with the header"


def test_version_and_null_argument_errors(lib):
    from slemm_amd import _lib
    assert lib.slemm_gpu_version() >= 1
    h = C.c_void_p()
    st = lib.slemm_gpu_geno_create(None, 10, 10, 3, None, 0, None, 0, C.byref(h))
    assert st == _lib.ERR_ARG
    assert b"null" in lib.slemm_gpu_last_error()
    st = lib.slemm_gpu_op(None, 0, None, 0.5, None, 1, 1, None, 1, 0, None)
    assert st == _lib.ERR_ARG
    st = lib.slemm_gpu_lseed_info(None, None, None, None)
    assert st == _lib.ERR_ARG


def test_product_package_does_not_import_the_oracle():
    for path in (ROOT / "slemm_amd").rglob("*.py"):
        src = path.read_text()
        assert not re.search(r"^\s*(import|from)\s+oracle\b", src, flags=re.M), path
