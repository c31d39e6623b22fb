"""MI355X-native indirect par_loop for the PyOP2 API (single rank).

``from pyop2_amd import op2`` gives the ``op2`` namespace (Set, Map, Dat,
Global, Mat, Sparsity, Kernel, par_loop, access descriptors).
"""
__version__ = "0.1.0"
