"""Launch-parameter autotuner of the hip backend.

For every loop signature (kernel, argument kinds and accesses, loop mode)
and mesh-size bucket (log2 of the iteration set size) it times candidate
launch parameters on the live loop and keeps the fastest: workgroup size
for direct loops; workgroup size and elements per thread (the owned range)
for owner-computes loops; partition size and workgroup size for coloured
loops.  Every candidate computes the same result (owner and direct loops
bit-identically: their sums run in a fixed element order), so tuning only
changes speed.  The Dats, Globals and Mats a loop writes are copied before
the trials and restored after them, so tuning leaves no trace in the data.

The table persists as JSON in the code-object cache directory
(``autotune.json``), keyed by signature and bucket.
"""

import hashlib
import json
import math
import os

from . import configuration as cfg

CANDIDATES = {
    'stream': [{'hip_block_size': b} for b in (128, 256, 512)],
    'owner': [{'hip_owner_block_size': b, 'hip_owner_fill': f}
              for b in (256, 512) for f in (0.6, 0.9, 1.2)],
    # Mat blocks: workgroup size, elements per thread, and the LDS a block
    # may use (0: a CU's share at 8 waves per SIMD; 2x that: fewer, larger
    # owned ranges with less redundant element work)
    'owner_mat': [{'hip_owner_mat_block_size': b, 'hip_owner_mat_fill': f,
                   'hip_owner_mat_lds': lds * b // 256}
                  for b in (128, 256, 512) for f in (0.4, 0.6, 0.7, 0.9)
                  for lds in (0, 40960)],
    # coloured plans: partition size, workgroup size and colouring depth
    # (two levels, blocks then threads as plan.pyx; or one level, elements)
    'coloured': [{'hip_partition_size': p, 'hip_block_size': b, 'hip_colouring': c}
                 for p in (128, 256, 512, 1024) for b in (128, 256) if b <= p
                 for c in ('blocks', 'elements')],
}

_table = None


def _path():
    from . import jit
    return os.path.join(jit.cache_dir(), 'autotune.json')


def table():
    global _table
    if _table is None:
        try:
            with open(_path()) as f:
                _table = json.load(f)
        except (IOError, ValueError):
            _table = {}
    return _table


def _save():
    path = _path()
    try:
        os.makedirs(os.path.dirname(path), exist_ok=True)
        tmp = path + '.%d.tmp' % os.getpid()
        with open(tmp, 'w') as f:
            json.dump(table(), f, indent=1, sort_keys=True)
        os.replace(tmp, path)
    except OSError:
        pass


def signature(pl, mode):
    """Table key of a par_loop: kernel, arguments, mode, size bucket."""
    k = pl.kernel
    h = hashlib.md5(k.code.encode()).hexdigest()[:12]
    args = []
    for a in pl.args:
        kind = ('mat' if a._is_mat else 'gbl' if a._is_global else
                'dir' if a._is_direct else 'ind')
        args.append('%s:%s:%s' % (kind, a.access._mode, getattr(a.data, 'cdim', 0)))
    n = max(1, pl._it_space.size)
    return '%s/%s/%s/%s/2^%d' % (k.name, h, mode, ','.join(args), int(math.log2(n)))


def params_for(pl, mode, trial):
    """Tuned parameters of ``pl`` (a dict of configuration entries).
    ``trial(params)`` runs the loop under ``params`` and returns seconds per
    call; it is used only when the signature has no entry yet."""
    key = signature(pl, mode)
    t = table()
    if key in t:
        return t[key]['params']
    best, best_s, tried = None, None, []
    kind = mode
    if mode == 'owner' and any(a._is_mat for a in pl.args):
        kind = 'owner_mat'
    for cand in CANDIDATES.get(kind, [{}]):
        try:
            sec = trial(cand)
        except Exception as e:   # a candidate that cannot run (LDS, plan) is skipped
            tried.append({'params': cand, 'error': type(e).__name__})
            continue
        tried.append({'params': cand, 'us': round(sec * 1e6, 3)})
        if best_s is None or sec < best_s:
            best, best_s = cand, sec
    if best is None:
        best = {}
    t[key] = {'params': best, 'us': round((best_s or 0.0) * 1e6, 3), 'tried': tried}
    _save()
    return best


def reset(persisted=False):
    """Forget the in-memory table (it reloads from its file on next use);
    with ``persisted`` delete the file too."""
    global _table
    _table = None
    if persisted:
        try:
            os.remove(_path())
        except OSError:
            pass
