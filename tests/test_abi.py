"""CPU-side checks of the C ABI library: it loads without a GPU and exports
every function include/pic_abi.h declares; calls without a device fail with
a status code instead of crashing."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "pic_abi.h")


def declared_functions():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^(?:int|void|const char\*)\s+(pic_\w+)\(", text, re.M)))


def test_header_declares_expected_surface():
    names = declared_functions()
    import pic_gpu

    assert names == sorted(pic_gpu.ABI_FUNCTIONS)


def test_header_constants_match_binding():
    """Every solver, tolerance, flag and field constant the header defines
    has the same value in the Python binding."""
    import pic_gpu

    text = open(HEADER).read()
    consts = re.findall(r"^#define PIC_((?:SOLVER|SOLVE|TOL|FIELD|SFIELD|DEPOSIT|SURF)_\w+)\s+(\d+)", text, re.M)
    assert len(consts) >= 20
    for name, value in consts:
        assert getattr(pic_gpu, name) == int(value), name
    kcount = int(re.search(r"^#define PIC_K_COUNT (\d+)", text, re.M).group(1))
    assert len(pic_gpu.K_NAMES) == kcount


def test_library_exports_every_declared_symbol():
    import pic_gpu

    if not os.path.exists(pic_gpu.LIB_PATH):
        pytest.fail("libpic_mi355.so not built (run __graft_entry__.build())")
    lib = ctypes.CDLL(pic_gpu.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_bad_arguments_return_status_codes():
    import pic_gpu

    L = pic_gpu.load_library()
    assert L.pic_create(None, None) == -1
    g = pic_gpu.GridDesc()
    g.ni, g.nj, g.nk = 2, 8, 8  # too few nodes
    h = ctypes.c_void_p()
    assert L.pic_create(ctypes.byref(g), ctypes.byref(h)) == -1
    assert L.pic_sync(None) == -1
    assert L.pic_push(None, 0, 1e-9) == -1
    assert L.pic_last_error(None) == b"null context"
