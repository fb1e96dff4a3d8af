"""Stand-in backend after ``op2.exit``: every constructor raises
(reference: pyop2/finalised.py:38-75)."""


def _raise(*args, **kwargs):
    raise RuntimeError("op2.exit has been called")


class _Finalised(object):
    def __init__(self, *args, **kwargs):
        _raise()


for _n in ('Access', 'IterationSpace', 'Set', 'Halo', 'Kernel', 'Dat', 'Mat',
           'Const', 'Global', 'Map', 'Sparsity', 'Solver', 'Arg', 'ParLoop'):
    globals()[_n] = type(_n, (_Finalised,), {})

par_loop = _raise
solve = _raise
