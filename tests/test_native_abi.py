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


def test_backend_refuses_to_run_without_the_library(tmp_path):
    """A copy of the package without its built library cannot import the
    hip backend: ImportError naming libop2hip.so (no silent fallback)."""
    import shutil
    import subprocess
    import sys
    pkg = tmp_path / 'pyop2_amd'
    shutil.copytree(os.path.join(ROOT, 'pyop2_amd'), str(pkg),
                    ignore=shutil.ignore_patterns('lib', '__pycache__', '*.so', '*.o'))
    code = ("import sys; sys.path.insert(0, %r)\n"
            "try:\n    import pyop2_amd.hip\nexcept ImportError as e:\n"
            "    sys.exit(0 if 'libop2hip.so' in str(e) else 2)\nsys.exit(1)\n" % str(tmp_path))
    r = subprocess.run([sys.executable, '-c', code], cwd=str(tmp_path))
    assert r.returncode == 0
