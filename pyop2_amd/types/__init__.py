from pyop2_amd.types.access import Access, READ, WRITE, RW, INC, MIN, MAX  # noqa: F401
from pyop2_amd.types.set import Set, Subset, SetPartition, MixedSet, ExtrudedSet  # noqa: F401
from pyop2_amd.types.dataset import DataSet  # noqa: F401
from pyop2_amd.types.map import Map, PermutedMap, ComposedMap  # noqa: F401
from pyop2_amd.types.mixed import MixedDataSet, MixedMap, MixedDat  # noqa: F401
from pyop2_amd.types.dat import AbstractDat, Dat, DatView  # noqa: F401
from pyop2_amd.types.glob import Global, Constant  # noqa: F401
from pyop2_amd.types.mat import Sparsity, Mat  # noqa: F401
from pyop2_amd.types.region import (IterationRegion, ON_BOTTOM, ON_TOP,  # noqa: F401
                                     ON_INTERIOR_FACETS, ALL)
