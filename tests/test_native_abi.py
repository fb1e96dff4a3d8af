"""The C-ABI library loads and exports every entry point include/op2hip.h
declares, with no compute call (runs without a GPU)."""

import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    src = open(os.path.join(ROOT, 'include', 'op2hip.h')).read()
    src = re.sub(r'/\*.*?\*/', '', src, flags=re.S)
    return sorted(set(re.findall(r'\b(op2hip_\w+)\s*\(', src)))


def test_header_declares_entry_points():
    names = declared()
    assert 'op2hip_launch' in names and 'op2hip_cgb_step' in names
    assert len(names) >= 25


def test_library_exports_every_declared_symbol():
    from pyop2_amd import _native
    lib = ctypes.CDLL(_native.LIB_PATH)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_matches_header():
    from pyop2_amd import _native
    assert set(_native.EXPORTED) == set(declared())


def test_compile_needs_no_gpu():
    """hiprtc compiles a generated loop on a host without a GPU."""
    from pyop2_amd import jit
    code = jit.compile_source('extern "C" __global__ void k(double* x) { x[threadIdx.x] = 1.0; }')
    assert len(code) > 1000


def test_cg_system_layout_matches_library():
    from pyop2_amd import _native, cg
    import ctypes as C
    assert C.sizeof(cg.CGSystem) == _native.lib.op2hip_cgb_sizeof_system()
