"""Stand-in backend before ``op2.init``: every constructor raises
(reference: pyop2/void.py:36-88)."""


def _raise(*args, **kwargs):
    raise RuntimeError("Please call op2.init to select a backend")


class _Void(object):
    def __init__(self, *args, **kwargs):
        _raise()


for _n in ('Access', 'IterationSpace', 'Set', 'Halo', 'Kernel', 'Dat', 'Mat',
           'Const', 'Global', 'Map', 'Sparsity', 'Solver', 'Arg', 'ParLoop'):
    globals()[_n] = type(_n, (_Void,), {})

par_loop = _raise
solve = _raise
