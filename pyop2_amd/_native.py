"""ctypes binding of ``libop2hip.so``, the C-ABI library of the hip backend.

Every entry point returns an int status (0 ok, <0 error) and the message of
the last failure is read with ``op2hip_last_error``; :func:`check` turns a
failing status into a :class:`NativeError`.  The library is built in-tree by
:mod:`pyop2_amd.build`; when it is missing this module raises on import, so
nothing on the hip path can run without the native code.
"""

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, 'lib', 'libop2hip.so')


class NativeError(RuntimeError):
    """A call into libop2hip failed."""


if not os.path.exists(LIB_PATH):
    raise ImportError("libop2hip.so is not built (expected at %s); run "
                      "`python -m pyop2_amd.build` first" % LIB_PATH)



def _preload_torch_hip():
    """Load the HIP runtime torch ships before libop2hip.so.

    torch brings its own libamdhip64/libhsa-runtime64/libhiprtc (same
    sonames as /opt/rocm's).  libop2hip.so must bind to those copies: two HIP
    runtimes in one process each open the GPU, and the second finds none.
    Loading torch's first makes the dynamic linker resolve libop2hip.so's
    dependencies to them by soname."""
    try:
        import torch
    except ImportError:
        return
    libdir = os.path.join(os.path.dirname(torch.__file__), 'lib')
    for name in ('libhsa-runtime64.so', 'libamdhip64.so', 'libamd_comgr.so', 'libhiprtc.so'):
        path = os.path.join(libdir, name)
        if os.path.exists(path):
            ctypes.CDLL(path, mode=ctypes.RTLD_GLOBAL)


_preload_torch_hip()
lib = ctypes.CDLL(LIB_PATH)

c_int = ctypes.c_int
c_ll = ctypes.c_longlong
c_size = ctypes.c_size_t
c_void = ctypes.c_void_p
c_double = ctypes.c_double
c_char_p = ctypes.c_char_p
P = ctypes.POINTER

_SIGS = {
    'op2hip_last_error': (c_char_p, []),
    'op2hip_init': (c_int, [c_int]),
    'op2hip_device_arch': (c_int, [c_int, c_char_p, c_int]),
    'op2hip_device_limits': (c_int, [c_int, P(c_int), P(c_int), P(c_int)]),
    'op2hip_compile': (c_int, [c_char_p, P(c_char_p), c_int, c_char_p, P(c_void), P(c_size),
                               P(c_int)]),
    'op2hip_free_host': (None, [c_void]),
    'op2hip_module_load': (c_int, [c_void, c_size, P(c_void)]),
    'op2hip_module_destroy': (c_int, [c_void]),
    'op2hip_module_function': (c_int, [c_void, c_char_p, P(c_void)]),
    'op2hip_module_global': (c_int, [c_void, c_char_p, P(c_void), P(c_size)]),
    'op2hip_set_global': (c_int, [c_void, c_void, c_size, c_void]),
    'op2hip_store_small': (c_int, [c_void, c_void, c_size, c_void]),
    'op2hip_launch': (c_int, [c_void, ctypes.c_uint, ctypes.c_uint, ctypes.c_uint, c_void,
                              P(ctypes.c_uint64), c_int]),
    'op2hip_function_attrs': (c_int, [c_void, P(c_int), P(c_int), P(c_int)]),
    'op2hip_event_create': (c_int, [P(c_void)]),
    'op2hip_event_destroy': (c_int, [c_void]),
    'op2hip_event_record': (c_int, [c_void, c_void]),
    'op2hip_event_elapsed_ms': (c_int, [c_void, c_void, P(ctypes.c_float)]),
    'op2hip_stream_sync': (c_int, [c_void]),
    'op2hip_plan_build': (c_int, [c_int, c_int, c_int, c_int, P(c_void), P(c_int), P(c_int),
                                  P(c_int), c_int, P(c_int), P(c_void)]),
    'op2hip_plan_scalar': (c_int, [c_void, c_int, P(ctypes.c_int64)]),
    'op2hip_plan_array': (c_int, [c_void, c_int, c_void]),
    'op2hip_plan_destroy': (c_int, [c_void]),
    'op2hip_plan_verify': (c_int, [c_int, c_int, c_void, c_void, c_void, c_void, c_void, c_int,
                                   P(c_void), P(c_int), P(c_int), P(c_int), c_int, P(c_int),
                                   P(c_int)]),
    'op2hip_plan_workspace': (c_ll, [c_int, c_int, c_int, P(c_int), c_int]),
    'op2hip_plan_dbuild': (c_int, [c_int, c_int, c_int, c_int, P(c_void), P(c_int), P(c_int),
                                   P(c_int), c_int, P(c_int), c_void, c_ll, c_void, P(c_void),
                                   P(c_int), P(c_int), c_void]),
    'op2hip_plan_dfill': (c_int, [c_void] * 9 + [c_void]),
    'op2hip_plan_ddiscard': (c_int, [c_void]),
    'op2hip_owner_plan_build': (c_int, [c_int, c_int, c_int, P(c_void), P(c_int), P(c_int), c_int,
                                        c_void, P(c_void)]),
    'op2hip_owner_plan_scalar': (c_int, [c_void, c_int, P(ctypes.c_int64)]),
    'op2hip_owner_plan_array': (c_int, [c_void, c_int, c_void]),
    'op2hip_owner_plan_destroy': (c_int, [c_void]),
    'op2hip_owner_plan_workspace': (c_ll, [c_int, c_int, c_int, c_int]),
    'op2hip_owner_plan_dbuild': (c_int, [c_int, c_int, c_int, P(c_void), P(c_int), P(c_int), c_int,
                                         c_void, c_void, c_ll, c_void, P(c_void), P(c_ll),
                                         c_void]),
    'op2hip_owner_plan_dfill': (c_int, [c_void] + [c_void] * 11 + [c_int, c_void]),
    'op2hip_owner_plan_ddiscard': (c_int, [c_void]),
    'op2hip_mat_program_workspace': (c_ll, [c_ll]),
    'op2hip_mat_program_device': (c_int, [c_int] + [c_void] * 8 + [c_int, c_int, c_void, c_int,
                                                                  c_int, c_int, c_void, c_void,
                                                                  c_int, c_int, c_ll, c_int,
                                                                  c_void, c_ll, c_void, c_void,
                                                                  c_void]),
    'op2hip_mat_rows_workspace': (c_ll, [c_int, c_int, c_int, c_int]),
    'op2hip_mat_rows_build': (c_int, [c_int, c_int, c_void, c_int, c_int, c_void, c_int, c_int,
                                      c_void, c_void, c_void, c_int, c_int, c_int, c_void, c_ll,
                                      c_void, c_void, c_void, c_void, c_void, P(c_ll), c_void]),
    'op2hip_mat_rows_pack16': (c_int, [c_void, c_ll, c_void, c_void]),
    'op2hip_stage_max_keys': (c_int, []),
    'op2hip_stage_lists': (c_int, [c_int, c_void, c_void, c_void, c_int, c_int, c_ll, c_void,
                                   c_void, c_void, c_void, c_void]),
    'op2hip_stage_compact': (c_int, [c_int, c_void, c_int, c_void, c_void, c_void, c_void,
                                     c_void]),
    'op2hip_mat_program': (c_int, [c_int, c_void, c_void, c_void, c_void, c_void, c_void, c_void,
                                   c_void, c_int,
                                   c_void, c_int, c_int, c_int, c_void, c_void, c_int, c_int,
                                   c_int, c_void, c_void]),
    'op2hip_order_bfs': (c_int, [c_int, c_int, c_void, c_int, c_void]),
    'op2hip_module_compile': (c_int, [ctypes.c_char_p, ctypes.c_char_p, c_void, c_int,
                                      ctypes.c_char_p, P(c_void), P(c_void)]),
    'op2hip_module_set_const': (c_int, [c_void, ctypes.c_char_p, c_void, c_size, c_void]),
    'op2hip_pcg': (c_int, [c_int, c_void, c_void, c_void, c_void, c_void, c_void, ctypes.c_double,
                           ctypes.c_double, ctypes.c_double, c_int, P(c_int),
                           P(ctypes.c_double), P(c_int), c_void]),
    'op2hip_pcg_noalloc': (c_int, [c_int, c_void, c_void, c_void, c_void, c_void, c_void,
                                   ctypes.c_double, ctypes.c_double, ctypes.c_double, c_int,
                                   P(c_int), P(ctypes.c_double), P(c_int), c_void]),
    'op2hip_pcg_noalloc_max_rows': (c_int, []),
    'op2hip_order_bisect': (c_int, [c_int, c_int, c_void, c_int, c_int, c_void]),
    'op2hip_order_first_touch': (c_int, [c_int, c_int, c_void, c_int, c_void, c_void]),
    'op2hip_sparsity_workspace': (c_ll, [c_int, c_int, P(c_ll), P(c_int), P(c_int)]),
    'op2hip_sparsity_build': (c_int, [c_int, c_int, c_int, c_int, P(c_void), P(c_int), P(c_void),
                                      P(c_int), P(c_ll), c_void, c_ll, c_void, P(c_void), P(c_ll),
                                      c_void]),
    'op2hip_sparsity_fill': (c_int, [c_void, c_void, c_void, c_void, c_void]),
    'op2hip_sparsity_discard': (c_int, [c_void]),
    'op2hip_spmv': (c_int, [c_int, c_ll, c_void, c_void, c_void, c_void, c_void, c_void]),
    'op2hip_csr_zero_rows': (c_int, [c_int, c_void, c_void, c_void, c_void, c_double, c_void]),
    'op2hip_cgb_create': (c_int, [c_int, c_ll, c_int, P(c_void)]),
    'op2hip_cgb_destroy': (c_int, [c_void]),
    'op2hip_cgb_load': (c_int, [c_void, c_int, c_void, c_void]),
    'op2hip_cgb_step': (c_int, [c_void, c_int, c_void]),
    'op2hip_cgb_step_graph': (c_int, [c_void, c_int, c_void]),
    'op2hip_cgb_read': (c_int, [c_void, c_void, c_void]),
    'op2hip_cgb_history': (c_int, [c_void, c_int, c_void, c_void]),
    'op2hip_cgb_sizeof_system': (c_int, []),
    'op2hip_cgb_chunk_rows': (c_int, []),
    'op2hip_krylov_work': (c_ll, [c_int, c_int, c_int]),
    'op2hip_krylov_solve': (c_int, [c_int, c_int, c_void, c_void, c_void, c_void, c_void, c_void,
                                    c_ll, c_int, c_int, c_double, c_double, c_double, c_int,
                                    c_void, c_int, P(c_int), P(c_double), P(c_int), c_void]),
}

for _name, (_res, _args) in _SIGS.items():
    _f = getattr(lib, _name)
    _f.restype = _res
    _f.argtypes = _args

EXPORTED = tuple(sorted(_SIGS))


def last_error():
    msg = lib.op2hip_last_error()
    return msg.decode(errors='replace') if msg else ''


def check(status, what=''):
    """Raise :class:`NativeError` for a failing status."""
    if status != 0:
        raise NativeError("%s failed (%d): %s" % (what or 'libop2hip call', status, last_error()))
    return status


def int_array(values):
    """A ctypes int array of ``values``."""
    return (c_int * len(values))(*values)


def ptr_array(values):
    """A ctypes void* array of integer addresses."""
    return (c_void * len(values))(*[c_void(v) for v in values])
