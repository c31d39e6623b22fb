"""The ``op2`` namespace (``pyop2/op2.py:43-68``): ``from pyop2_amd import op2``."""
from pyop2_amd.configuration import configuration
from pyop2_amd.exceptions import *  # noqa: F401,F403
from pyop2_amd.local_kernel import CStringLocalKernel, Kernel, LocalKernel  # noqa: F401
from pyop2_amd.global_kernel import (GlobalKernel, GlobalKernelArg, DatKernelArg,  # noqa: F401
                                     MatKernelArg, MapKernelArg)
from pyop2_amd.parloop import (GlobalParloopArg, DatParloopArg, MatParloopArg,  # noqa: F401
                               Parloop, ParLoop, par_loop, parloop, PassthroughArg)
from pyop2_amd.types import (Access, READ, WRITE, RW, INC, MIN, MAX,  # noqa: F401
                             Set, Subset, SetPartition, MixedSet, ExtrudedSet, DataSet, Map, Dat, DatView,
                             Global, Constant, Sparsity, Mat, PermutedMap, ComposedMap,
                             MixedDataSet, MixedMap, MixedDat,
                             IterationRegion, ON_BOTTOM, ON_TOP, ON_INTERIOR_FACETS, ALL)

__all__ = ['configuration', 'READ', 'WRITE', 'RW', 'INC', 'MIN', 'MAX',
           'Set', 'Subset', 'ExtrudedSet', 'DataSet', 'MixedSet', 'MixedDataSet', 'Map', 'MixedMap',
           'PermutedMap', 'ComposedMap', 'MixedDat',
           'ON_BOTTOM', 'ON_TOP', 'ON_INTERIOR_FACETS', 'ALL',
           'Dat', 'DatView', 'Global', 'Constant',
           'Kernel', 'Sparsity', 'Mat', 'par_loop', 'parloop', 'ParLoop', 'Parloop',
           'GlobalKernel', 'DatKernelArg', 'MatKernelArg', 'GlobalKernelArg', 'MapKernelArg',
           'GlobalParloopArg', 'DatParloopArg', 'MatParloopArg', 'PassthroughArg',
           'init', 'exit']

_initialised = False


def initialised():
    return _initialised


def init(**kwargs):
    """Reconfigure (``pyop2/op2.py:82-108``).  The device is selected lazily
    on the first par_loop (``device=N`` picks the HIP ordinal)."""
    global _initialised
    configuration.reconfigure(**kwargs)
    _initialised = True


def exit():
    """Reset the configuration."""
    global _initialised
    configuration.reset()
    _initialised = False
