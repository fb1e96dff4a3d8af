"""Configuration: a module that behaves as a dictionary of options.

Lookup order (reference: pyop2/configuration.py:54-106): keyword arguments to
:func:`configure`, then the file named by ``config=``, else ``./pyop2.yaml``
when present, then the defaults below.  Options are read as ``cfg['key']``
or ``cfg.key``.
"""

import sys
import types
import warnings

import yaml

DEFAULTS = {
    'log_level': 'WARN',
    'python_plan': True,
    'backend': 'hip',
    'debug': 0,
    'dump-gencode': False,
    'dump-gencode-path': '/tmp/%(kernel)s-%(time)s.hip',
    # hip backend options
    'hip_partition_size': 0,      # 0: taken from the autotuner table
    'hip_block_size': 0,          # 0: taken from the autotuner table
    'hip_reorder': True,          # locality-ordered partitions for indirect loops
    'hip_colouring': 'auto',      # coloured loops: 'elements' (one colour per element,
                                  # blocks of one colour), 'blocks' (block and thread
                                  # colours) or 'auto' (elements for Mat assembly)
    'hip_autotune': False,        # tune launch parameters per loop and mesh-size bucket
    'hip_stream_blocks': 4096,    # most workgroups of a direct loop (more rows: grid-stride)
    'hip_stream_unroll': 1,       # elements per thread and chunk of direct loops (1 or 2)
    'hip_owner_fill': 1.0,        # owner blocks run about this many elements per thread
    'hip_owner_block_size': 512,  # workgroup size of owner-computes loops (0: hip_block_size)
    'hip_owner_mat_block_size': 256,  # ... of owner-computes Mat assembly
    'hip_owner_mat_fill': 0.7,    # elements per thread of owner-computes Mat assembly
    'hip_stage_reads': False,     # stage read-only multi-component direct Dats through LDS
    'hip_stage_writes': True,     # stage written ones through LDS (else whole-row vector stores)
    'hip_owner_stage': False,     # owner blocks stage their owned rows of gathered Dats in LDS
    'hip_owner_prefetch': 1,      # owner write-back loads issued before the elements run:
                                  # 0 none, 1 the first target's lookups, 2 also its old value
    'hip_owner_mat_lds': 0,       # LDS bytes an owner Mat block may use (0: twice a CU's
                                  # share at full occupancy, 40 KB for 256 threads)
    'hip_mat_rows': True,         # owner Mat assembly by row blocks (CSR slice in LDS)
                                  # for local matrices of at most 16 values
    'hip_mat_rows_block_size': 64,  # workgroup size of row blocks (one wavefront)
    'hip_mat_rows_k': 64,         # row nodes per row block (at most the workgroup size)
    'hip_mat_rows_lds': 0,        # LDS bytes a row block may use (0: 52 KB)
    'hip_mat_rows_stage': False,  # row blocks stage their elements' rows in LDS
    'hip_defer_reductions': True,  # an FP64 Global's combine pass rides on the next launch
    'hip_owner_waves': 0,         # waves per SIMD owner loops are compiled for (0: compiler's choice)
    'hip_device_plan': True,      # owner plans built on the device (False: the host build)
    'hip_layout': True,           # locality renumbering of device storage
    'hip_layout_order': 'bisect',  # adjacency order of sets no map orders: 'bisect' or 'bfs'
    'hip_cache_dir': None,        # on-disk code-object cache (None: ~/.cache/pyop2_amd)
}


_MISSING = object()


class ConfigModule(types.ModuleType):
    """Module object answering ``cfg[key]`` and ``cfg.key``."""

    _config = None
    # bumped on every change, so callers can cache values derived from the
    # configuration
    _version = 0

    def version(self):
        return self._version

    def configure(self, **kwargs):
        entries = dict(DEFAULTS)
        cfgfile = kwargs.pop('config', None)
        if cfgfile is not None:
            try:
                src = cfgfile.read() if hasattr(cfgfile, 'read') else open(cfgfile).read()
                entries.update(yaml.safe_load(src) or {})
            except IOError:
                pass
        else:
            try:
                with open('pyop2.yaml') as f:
                    entries.update(yaml.safe_load(f) or {})
            except IOError:
                pass
        entries.update(kwargs)
        self._config = entries
        self._version += 1
        if entries.get('debug', 0) > 0:
            warnings.simplefilter("always")

    def reset(self):
        """Forget all configuration entries."""
        self._config = None
        self._version += 1

    def override(self, **kwargs):
        """Context manager setting entries for the duration of a block."""
        return _Override(self, kwargs)

    def get(self, key, default=None):
        """Return an option, falling back to the defaults before configure()."""
        if self._config is None:
            return DEFAULTS.get(key, default)
        return self._config.get(key, DEFAULTS.get(key, default))

    def __getitem__(self, key):
        if not self._config:
            raise KeyError(key)
        return self._config[key]

    def __getattr__(self, name):
        if name.startswith('__') or not self._config:
            raise AttributeError(name)
        try:
            return self._config[name]
        except KeyError:
            raise AttributeError(name)


class _Override(object):
    """Entries of a configuration set for the duration of a block."""

    def __init__(self, mod, kwargs):
        self.mod, self.kwargs = mod, kwargs

    def __enter__(self):
        mod, kwargs = self.mod, self.kwargs
        if mod._config is None:
            mod._config = dict(DEFAULTS)
        self.saved = {k: mod._config.get(k, _MISSING) for k in kwargs}
        mod._config.update(kwargs)
        if kwargs:
            mod._version += 1

    def __exit__(self, *exc):
        mod = self.mod
        for k, v in self.saved.items():
            if v is _MISSING:
                mod._config.pop(k, None)
            else:
                mod._config[k] = v
        if self.saved:
            mod._version += 1
        return False


sys.modules[__name__].__class__ = ConfigModule
