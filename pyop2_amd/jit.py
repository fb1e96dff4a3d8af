"""hiprtc compilation of generated loop sources and loaded modules.

Compilation goes through ``op2hip_compile`` (hiprtc for gfx950, with an
on-disk code-object cache keyed by a hash of source, options and hiprtc
version) and ``op2hip_module_load``.  :func:`compile_source` needs no GPU;
:class:`Module` does.
"""

import ctypes
import os

from . import _native as nat
from . import configuration as cfg

ARCH = os.environ.get('OP2HIP_ARCH', 'gfx950')

# Loop kernels are compiled without FP contraction, so FP64 results follow
# the IEEE operation order of the kernel source (as the host oracle does),
# and without packed FP32 instructions.
HIPRTC_OPTS = ['--offload-arch=%s' % ARCH, '-O3', '-std=c++17', '-ffp-contract=off',
               '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops']


def cache_dir():
    d = cfg.get('hip_cache_dir')
    if d:
        return d
    return os.path.join(os.path.expanduser('~'), '.cache', 'pyop2_amd', 'co')


def compile_source(src, opts=None):
    """Compile HIP source to a gfx950 code object (bytes)."""
    opts = list(HIPRTC_OPTS if opts is None else opts)
    copts = (ctypes.c_char_p * len(opts))(*[o.encode() for o in opts])
    code = ctypes.c_void_p()
    size = ctypes.c_size_t()
    hit = ctypes.c_int()
    nat.check(nat.lib.op2hip_compile(src.encode(), copts, len(opts), cache_dir().encode(),
                                     ctypes.byref(code), ctypes.byref(size), ctypes.byref(hit)),
              'hiprtc compile')
    try:
        return ctypes.string_at(code.value, size.value)
    finally:
        nat.lib.op2hip_free_host(code)


class Module(object):
    """A loaded code object with its kernels and global symbols."""

    def __init__(self, code):
        self._code = code
        h = ctypes.c_void_p()
        buf = ctypes.create_string_buffer(code, len(code))
        nat.check(nat.lib.op2hip_module_load(buf, len(code), ctypes.byref(h)), 'module load')
        self._h = h
        self._fns = {}
        self._globals = {}

    def function(self, name):
        f = self._fns.get(name)
        if f is None:
            fn = ctypes.c_void_p()
            nat.check(nat.lib.op2hip_module_function(self._h, name.encode(), ctypes.byref(fn)),
                      'module function %s' % name)
            f = self._fns[name] = fn
        return f

    def global_symbol(self, name):
        g = self._globals.get(name)
        if g is None:
            p = ctypes.c_void_p()
            n = ctypes.c_size_t()
            nat.check(nat.lib.op2hip_module_global(self._h, name.encode(), ctypes.byref(p),
                                                   ctypes.byref(n)), 'module global %s' % name)
            g = self._globals[name] = (p, n.value)
        return g

    def __del__(self):
        try:
            if self._h:
                nat.lib.op2hip_module_destroy(self._h)
        except Exception:
            pass


def launch(fn, grid, block, stream, argv, lds=0):
    """Launch ``fn`` with 8-byte parameters ``argv`` (ints)."""
    n = len(argv)
    arr = (ctypes.c_uint64 * n)(*[int(v) & 0xFFFFFFFFFFFFFFFF for v in argv])
    nat.check(nat.lib.op2hip_launch(fn, int(grid), int(block), int(lds), ctypes.c_void_p(stream),
                                    arr, n), 'kernel launch')
