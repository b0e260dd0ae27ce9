"""CPU tests of the drop-in boundary: libmvbt.so loads without touching the GPU
and exports exactly the functions include/mvbt.h declares; the Python mirror
binds every one of them; header constants match the mirror."""
from __future__ import annotations

import ctypes
import os
import re
import subprocess

import mvgpubtree_amd as mv

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "mvbt.h")


def declared_functions():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:mvbt_status|const char\*)\s+(mvbt_\w+)\s*\(", text, re.M)))


def test_header_declares_the_mirrored_symbols():
    assert declared_functions() == sorted(mv.EXPORTED_SYMBOLS)


def test_library_loads_and_exports_every_declared_symbol():
    lib = mv.load_library()
    for name in declared_functions():
        assert hasattr(lib, name), name
        assert ctypes.cast(getattr(lib, name), ctypes.c_void_p).value


def test_exported_dynamic_symbols_match_header():
    out = subprocess.run(["nm", "-D", "--defined-only", mv.library_path()], capture_output=True,
                         text=True, check=True).stdout
    exported = sorted({l.split()[-1] for l in out.splitlines() if " T mvbt_" in l})
    assert exported == declared_functions()


def test_header_constants_match_mirror():
    text = open(HEADER).read()
    consts = dict(re.findall(r"#define\s+(MVBT_\w+)\s+(0x[0-9A-Fa-f]+)u", text))
    assert int(consts["MVBT_NOT_FOUND"], 16) == mv.NOT_FOUND
    assert int(consts["MVBT_LATEST"], 16) == mv.LATEST
    assert int(consts["MVBT_CONCURRENT"], 16) == mv.CONCURRENT
    assert int(consts["MVBT_FLAG_POOL_EXHAUSTED"], 16) == mv.FLAG_POOL_EXHAUSTED
    assert int(consts["MVBT_FLAG_RETIRE_OVERFLOW"], 16) == mv.FLAG_RETIRE_OVERFLOW
    assert int(consts["MVBT_FLAG_INVALID_KEY"], 16) == mv.FLAG_INVALID_KEY
    assert int(consts["MVBT_FLAG_VERSION_GAP"], 16) == mv.FLAG_VERSION_GAP
    assert int(consts["MVBT_FLAG_RANGE_OVERFLOW"], 16) == mv.FLAG_RANGE_OVERFLOW
    assert int(consts["MVBT_FLAG_CORRUPT"], 16) == mv.FLAG_CORRUPT
    assert int(consts["MVBT_FLAG_SNAPSHOT_ID"], 16) == mv.FLAG_SNAPSHOT_ID
    assert int(consts["MVBT_FLAG_STAMP_ORDER"], 16) == mv.FLAG_STAMP_ORDER
    assert int(consts["MVBT_FLAG_BAD_BATCH"], 16) == mv.FLAG_BAD_BATCH


def test_device_error_bits_match_header_flags():
    """The kernels raise kErr* bits (csrc/mvbt_common.hpp); the ABI documents
    them as MVBT_FLAG_* (include/mvbt.h): the two lists must agree bit for bit."""
    text = open(HEADER).read()
    flags = {k: int(v, 16) for k, v in re.findall(r"#define\s+MVBT_FLAG_(\w+)\s+(0x[0-9A-Fa-f]+)u", text)}
    common = open(os.path.join(ROOT, "mvgpubtree_amd", "csrc", "mvbt_common.hpp")).read()
    bits = {k: 1 << int(v) for k, v in re.findall(r"kErr(\w+)\s*=\s*1u\s*<<\s*(\d+)", common)}
    snake = {re.sub(r"(?<!^)(?=[A-Z])", "_", k).upper(): v for k, v in bits.items()}
    assert snake == flags


def test_null_handle_is_an_error_not_a_crash():
    lib = mv.load_library()
    assert lib.mvbt_insert(None, None, None, 1, None, 0) == 1
    assert lib.mvbt_release_snapshot(None, 0) == 1
    assert b"null" in lib.mvbt_last_error()
    assert lib.mvbt_create(None, None) == 1


def test_invalid_config_is_rejected_without_device_work():
    lib = mv.load_library()
    h = ctypes.c_void_p()
    cfg = mv._Config(kind=7, node_width=16, pool_nodes=1024, retire_capacity=1024, device=0)
    assert lib.mvbt_create(ctypes.byref(cfg), ctypes.byref(h)) == 1
    cfg = mv._Config(kind=1, node_width=24, pool_nodes=1024, retire_capacity=1024, device=0)
    assert lib.mvbt_create(ctypes.byref(cfg), ctypes.byref(h)) == 1
    cfg = mv._Config(kind=1, node_width=16, pool_nodes=1024, retire_capacity=1000, device=0)
    assert lib.mvbt_create(ctypes.byref(cfg), ctypes.byref(h)) == 1
    cfg = mv._Config(kind=1, node_width=16, pool_nodes=1 << 30, retire_capacity=1024, device=0)
    assert lib.mvbt_create(ctypes.byref(cfg), ctypes.byref(h)) == 1
    # node indices are 29-bit (the metadata word's top bits are flags)
    cfg = mv._Config(kind=1, node_width=16, pool_nodes=(1 << 29) - 1, retire_capacity=1024, device=0)
    assert lib.mvbt_create(ctypes.byref(cfg), ctypes.byref(h)) == 1
