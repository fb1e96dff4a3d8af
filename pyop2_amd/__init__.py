"""MI355X-native implementation of the PyOP2 programming model.

``from pyop2_amd import op2; op2.init(backend='hip')`` selects the GPU
backend; the API is that of PyOP2's ``op2`` module."""

__version__ = '0.1'
