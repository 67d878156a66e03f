"""The C-ABI library loads and exports every entry point include/fbme_hip.h
declares; planner queries work without a GPU (no compute calls)."""

import ctypes
import os
import re

import pytest

from conftest import ROOT
from fbme_amd import _lib


def _declared():
    text = open(os.path.join(ROOT, "include", "fbme_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b(fbme_[a-z0-9_]+)\s*\(", text)
    return sorted(set(names))


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    declared = _declared()
    assert len(declared) >= 15
    for name in declared:
        assert hasattr(lib, name), name
    assert set(declared) <= set(_lib.EXPORTED) | {"fbme_plan_create"}


def test_abi_version():
    assert _lib.load().fbme_abi_version() == _lib.ABI_VERSION


def test_dense_workspace_rejects_bad_shapes():
    """fbme_dense_workspace_bytes is host-only: 0 for every shape the dense
    posterior rejects, a positive size otherwise."""
    lib = _lib.load()
    assert lib.fbme_dense_workspace_bytes(2, 32, 32, 24, 24, 576) > 0
    assert lib.fbme_dense_workspace_bytes(1, 40, 30, 20, 15, 200) > 0
    for args in [(0, 32, 32, 24, 24, 576),        # no scenes
                 (70000, 32, 32, 24, 24, 576),    # beyond grid.y
                 (1, 32, 32, 40, 24, 576),        # M1 > N1
                 (1, 32, 32, 24, 24, 577),        # more observed than samples
                 (1, 64, 64, 48, 48, 2304),       # observed > 2048
                 (1, 32, 32, 24, 24, 0),
                 (1, -32, 32, 24, 24, 576)]:
        assert lib.fbme_dense_workspace_bytes(*args) == 0, args


def test_radix_codes_and_toeplitz_tiles():
    """fbme_radix_code and fbme_toeplitz_tiles are host-only.  A code exists
    exactly for the 5-smooth lengths up to 2048 and its nibbles (radices 2,
    3, 4, 5, 8, first stage lowest) multiply back to the length; the BTTB
    tile caps keep every LDS tile within the 16- (row) and 24-element
    (column) per-thread instances."""
    lib = _lib.load()

    def smooth(n):
        for p in (2, 3, 5):
            while n % p == 0:
                n //= p
        return n == 1

    for n in range(1, 2100):
        code = lib.fbme_radix_code(n)
        if n > 2048 or not smooth(n):
            assert code == 0xFFFFFFFF, n
            continue
        prod, c = 1, code
        while c:
            assert c & 15 in (2, 3, 4, 5, 8), (n, hex(code))
            prod *= c & 15
            c >>= 4
        assert prod == n, (n, hex(code))
    rmax, cmax = ctypes.c_int(), ctypes.c_int()
    for e1, e2 in [(8, 8), (1024, 1024), (1080, 12), (12, 2000), (2048, 3), (2048, 2048)]:
        assert lib.fbme_toeplitz_tiles(e1, e2, ctypes.byref(rmax), ctypes.byref(cmax)) == 0
        assert rmax.value >= 1 and rmax.value * (e2 + 1) <= 16 * 256
        assert 1 <= cmax.value <= min(e2, 16) and cmax.value * (e1 + 1) <= 24 * 256
    for e1, e2 in [(0, 8), (2049, 8), (8, 2049)]:
        assert lib.fbme_toeplitz_tiles(e1, e2, ctypes.byref(rmax), ctypes.byref(cmax)) != 0


def _plan(shapes, stable=False, **kw):
    lib = _lib.load()
    opt = _lib.default_options(stable)
    for k, v in kw.items():
        setattr(opt, k, v)
    flat = [v for s in shapes for v in s]
    arr = (ctypes.c_int * len(flat))(*flat)
    err = ctypes.create_string_buffer(256)
    h = lib.fbme_plan_create(len(shapes), arr, ctypes.byref(opt), err, 256)
    return h, err.value.decode()


def test_planner_budget_and_layout():
    lib = _lib.load()
    h, err = _plan([(256, 256, 160, 160), (64, 64, 30, 40), (192, 192, 60, 60)])
    assert h, err
    try:
        info = (ctypes.c_longlong * 8)()
        budgets = []
        for i in range(3):
            assert lib.fbme_plan_scene_info(h, i, info) == 0
            budgets.append(int(info[2]))
        # T_u = max(3, ceil(0.8 N / M)) (solver.py:1287-1309)
        assert budgets == [3, 3, 9]
        assert lib.fbme_plan_query(h, _lib.Q_MAX_ITERS) == 9
        assert lib.fbme_plan_query(h, _lib.Q_TOTAL_GRID) == 65536 + 4096 + 36864
        assert lib.fbme_plan_query(h, _lib.Q_TOTAL_SAMPLES) == 25600 + 1200 + 3600
        assert lib.fbme_plan_query(h, _lib.Q_WORKSPACE_BYTES) > 0
    finally:
        lib.fbme_plan_destroy(h)


@pytest.mark.parametrize("shape,msg", [
    ((32, 32, 40, 24), "0 < M <= N"),
    ((22, 32, 10, 10), "prime factor"),
    ((2048, 32, 10, 10), "outside"),
])
def test_planner_rejects_bad_shapes(shape, msg):
    h, err = _plan([shape])
    assert not h
    assert msg in err


def test_planner_rejects_bad_options():
    h, err = _plan([(32, 32, 24, 24)], polish_size=40)
    assert not h and "polish_size" in err
    h, err = _plan([(32, 32, 24, 24)], stable=True, polish_iterations=1)
    assert not h and "two polish" in err
