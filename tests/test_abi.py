"""C ABI checks that need no GPU: the library loads, exports every symbol that
include/cuSZp.h declares, keeps the reference enum values, and rejects bad
arguments before touching the device."""
import ctypes
import os
import re

import pytest

from cuszp_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(ROOT, "include", "cuSZp.h")).read()
    c_part = src.split("#ifdef __cplusplus\n}")[0]
    return sorted(set(re.findall(r"\b(cuSZp_\w+)\s*\(", c_part)))


def test_header_declares_reference_entries():
    names = _declared()
    assert "cuSZp_compress" in names and "cuSZp_decompress" in names
    per_dim = [n for n in names if re.match(r"cuSZp_(de)?compress_[123]D_(fixed|plain|outlier)_f(32|64)$", n)]
    assert len(per_dim) == 36


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in _declared() if not hasattr(lib, n)]
    assert not missing, missing
    assert sorted(_declared()) == sorted(_lib.EXPORTED)


def test_enum_values_match_reference():
    src = open(os.path.join(ROOT, "include", "cuSZp.h")).read()
    for name, val in [("CUSZP_DIM_1D", 1), ("CUSZP_DIM_2D", 2), ("CUSZP_DIM_3D", 3), ("CUSZP_MODE_FIXED", 0),
                      ("CUSZP_MODE_PLAIN", 1), ("CUSZP_MODE_OUTLIER", 2), ("CUSZP_TYPE_FLOAT", 0),
                      ("CUSZP_TYPE_DOUBLE", 1)]:
        assert re.search(rf"{name}\s*=\s*{val}\b", src), name


def test_dims_struct_layout():
    assert ctypes.sizeof(_lib.Uint3) == 12


@pytest.fixture(scope="module")
def L():
    return _lib.load()


def test_max_compressed_size(L, oracle):
    U = _lib.Uint3
    for n in (1, 31, 1000, 32768, 32769, 10**6):
        assert L.cuSZp_max_compressed_size(n, 1, U(0, 0, 0), 0, 1) == oracle.max_size(1, n)
    assert L.cuSZp_max_compressed_size(12 * 10 * 3, 2, U(12, 10, 3), 1, 2) == oracle.max_size(2, 360, (12, 10, 3))
    assert L.cuSZp_max_compressed_size(8 * 9 * 5, 3, U(8, 9, 5), 0, 0) == oracle.max_size(3, 360, (8, 9, 5))
    assert L.cuSZp_max_compressed_size(0, 1, U(0, 0, 0), 0, 1) == 0
    assert L.cuSZp_max_compressed_size(10, 2, U(5, 2, 1), 0, 1) == 0   # x % 4 != 0


@pytest.mark.parametrize(
    "args,status",
    [
        (dict(mode=3), 1),
        (dict(typ=2), 1),
        (dict(dim=4), 1),
        (dict(eb=0.0), 1),
        (dict(eb=-1.0), 1),
        (dict(eb=float("nan")), 1),
        (dict(eb=1e-50), 1),          # rounds to 0 in float
        (dict(dim=2, dims=(5, 2, 1), n=10), 2),
        (dict(dim=3, dims=(4, 0, 1), n=0), 2),
        (dict(dim=2, dims=(4, 4, 4), n=60), 2),
        (dict(cap=10), 4),
    ],
)
def test_compress_ex_rejects(L, args, status):
    a = dict(n=1000, eb=1e-3, dim=1, dims=(0, 0, 0), typ=0, mode=1, cap=1 << 20)
    a.update(args)
    size = ctypes.c_size_t(123)
    # non-null dummy pointers: validation must fail before any device access
    rc = L.cuSZp_compress_ex(ctypes.c_void_p(16), ctypes.c_void_p(16), a["cap"], a["n"], ctypes.byref(size),
                             a["eb"], a["dim"], _lib.Uint3(*a["dims"]), a["typ"], a["mode"], None)
    assert rc == status
    assert size.value == 0


def test_decompress_ex_rejects_short_stream(L):
    rc = L.cuSZp_decompress_ex(ctypes.c_void_p(16), ctypes.c_void_p(16), 100_000, 100, 1e-3, 1,
                               _lib.Uint3(0, 0, 0), 0, 1, None)
    assert rc == 5   # header alone is 4096 bytes


def test_zero_elements_is_a_noop(L):
    size = ctypes.c_size_t(9)
    assert L.cuSZp_compress_ex(None, None, 0, 0, ctypes.byref(size), 1e-3, 1, _lib.Uint3(0, 0, 0), 0, 1, None) == 0
    assert size.value == 0


def test_status_strings(L):
    for s in range(8):
        assert L.cuSZp_status_string(s)
    assert b"gfx950" in L.cuSZp_version()


def test_batch_item_layout():
    # cuszp_batch_item_t: 5 eight-byte fields, dim, dims{x,y,z}, type, mode
    B = _lib.BatchItem
    assert ctypes.sizeof(B) == 64
    assert [B.dim.offset, B.dims.offset, B.type.offset, B.mode.offset] == [40, 44, 56, 60]
    src = open(os.path.join(ROOT, "include", "cuSZp.h")).read()
    for name, val in [("CUSZP_BATCH_COMPRESS", 0), ("CUSZP_BATCH_DECOMPRESS", 1), ("CUSZP_BATCH_GRAPH", 1),
                      ("CUSZP_BATCH_TUNE", 2)]:
        assert re.search(rf"{name}\s*=\s*{val}\b", src), name


def test_batch_plan_rejects_before_device_work(L):
    # validation runs before any allocation or launch: no GPU needed
    U = _lib.Uint3
    good = _lib.BatchItem(16, 16, 1000, 1 << 20, 1e-3, 1, U(0, 0, 0), 0, 1)
    bad_eb = _lib.BatchItem(16, 16, 1000, 1 << 20, -1.0, 1, U(0, 0, 0), 0, 1)
    bad_dims = _lib.BatchItem(16, 16, 10, 1 << 20, 1e-3, 2, U(5, 2, 1), 0, 1)
    small = _lib.BatchItem(16, 16, 100_000, 10, 1e-3, 1, U(0, 0, 0), 0, 1)
    for items, status, index in [([good, bad_eb], 1, 1), ([bad_dims], 2, 0), ([good, good, small], 4, 2)]:
        arr = (_lib.BatchItem * len(items))(*items)
        plan = ctypes.c_void_p(0)
        bad = ctypes.c_size_t(99)
        rc = L.cuSZp_batch_plan_create(ctypes.byref(plan), arr, len(items), 0, 0, None, ctypes.byref(bad))
        assert rc == status and bad.value == index and not plan.value
    # decompress: a stream shorter than its header is corrupt
    arr = (_lib.BatchItem * 1)(small)
    plan = ctypes.c_void_p(0)
    assert L.cuSZp_batch_plan_create(ctypes.byref(plan), arr, 1, 1, 0, None, None) == 5
    assert L.cuSZp_batch_plan_create(ctypes.byref(plan), arr, 1, 7, 0, None, None) == 1   # bad direction


def test_server_struct_layouts():
    C, R = _lib.Completion, _lib.Request
    assert ctypes.sizeof(C) == 40
    assert [C.status.offset, C.direction.offset, C.offset.offset, C.d_out.offset, C.bytes.offset] == [8, 12, 16, 24, 32]
    assert ctypes.sizeof(R) == 80 and R.item.offset == 8 and R.arrival_s.offset == 72


def test_server_rejects_before_device_work(L):
    # argument checks run before any HIP call: no GPU needed
    s = ctypes.c_void_p(0)
    assert L.cuSZp_server_create(ctypes.byref(s), None, 1 << 20, 2, 16) == 1        # no pool
    assert L.cuSZp_server_create(ctypes.byref(s), 4096, 1 << 20, 0, 16) == 1        # no lanes
    assert L.cuSZp_server_create(ctypes.byref(s), 4096, 1 << 20, 5, 16) == 1        # too many lanes
    assert L.cuSZp_server_create(ctypes.byref(s), 4096, 1 << 20, 2, 0) == 1         # empty batches
    assert L.cuSZp_server_create(ctypes.byref(s), 4096, 100, 2, 16) == 1            # pool below one unit
    assert not s.value
    rid = ctypes.c_ulonglong(0)
    it = _lib.BatchItem(16, 16, 1000, 0, 1e-3, 1, _lib.Uint3(0, 0, 0), 0, 1)
    assert L.cuSZp_server_submit(None, 0, ctypes.byref(it), ctypes.byref(rid)) == 1
    n = ctypes.c_size_t(0)
    assert L.cuSZp_server_poll(None, None, 0, ctypes.byref(n), 0) == 1
    assert L.cuSZp_server_release(None, 1) == 1
