"""The C-ABI library loads and exports every symbol include/lym.h declares
(no compute calls: no GPU needed)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.join(os.path.dirname(__file__), "..")
HEADER = os.path.join(ROOT, "include", "lym.h")
LIB = os.path.join(ROOT, "tfleonardym_amd", "liblym.so")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(lym_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    for s in ["lym_wilson_apply", "lym_wilson_apply_square", "lym_dot", "lym_cgm", "lym_mr_multishift"]:
        assert s in syms


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(LIB):
        pytest.fail("liblym.so is not built; run `python -m tfleonardym_amd.build`")
    return ctypes.CDLL(LIB)


def test_exports_every_declared_symbol(lib):
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_python_binding_covers_header(lib):
    from tfleonardym_amd import _lib
    assert sorted(_lib.SIGNATURES) == declared_symbols()


def test_abi_version(lib):
    from tfleonardym_amd import _lib
    want = int(re.search(r"#define LYM_ABI_VERSION (\d+)", open(HEADER).read()).group(1))
    lib.lym_abi_version.restype = ctypes.c_int
    assert lib.lym_abi_version() == want == _lib.ABI_VERSION


def test_struct_layout_matches_header():
    """ctypes mirror of lym_wilson / lym_solve_stats has the C sizes."""
    from tfleonardym_amd import _lib
    # 4 int32 + 4 int32 dims + 3 double + 4 pointers + link_format + reserved
    assert ctypes.sizeof(_lib.LymWilson) == 4 * 4 + 4 * 4 + 3 * 8 + 4 * 8 + 2 * 4
    assert ctypes.sizeof(_lib.LymTileChoice) == 8 * 4 + 8
    assert ctypes.sizeof(_lib.LymSolveStats) == 4 * 4 + 8 + 8 + 64 * 4 + 64 * 8 + 64 * 8 + 64 * 4 + 2 * 4


def test_links_f32_bytes(lib):
    """Packed complex64 links: ceil(nc^2 / 2) 16-byte units per direction and
    site; unsupported shapes give 0."""
    lib.lym_links_f32_bytes.restype = ctypes.c_size_t
    lib.lym_links_f32_bytes.argtypes = [ctypes.c_int32, ctypes.c_int32, ctypes.c_int64]
    assert lib.lym_links_f32_bytes(4, 3, 1000) == 4 * 5 * 1000 * 16
    assert lib.lym_links_f32_bytes(4, 2, 1000) == 4 * 2 * 1000 * 16
    assert lib.lym_links_f32_bytes(2, 3, 384) == 2 * 5 * 384 * 16
    assert lib.lym_links_f32_bytes(3, 3, 10) == 0
    assert lib.lym_links_f32_bytes(4, 4, 10) == 0
    assert lib.lym_links_f32_bytes(4, 3, 0) == 0
