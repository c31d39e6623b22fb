"""Iteration regions of an extruded loop (``pyop2/types/__init__.py:14-31``).

A loop over an ExtrudedSet runs, per column with cells ``[bottom, cend)``
(``cend = layers[1] - 1``), the layers (``codegen/builder.py:771-795``):

- ``ALL``: every cell, ``[bottom, cend)``;
- ``ON_BOTTOM``: the bottom cell, ``[bottom, bottom + 1)``;
- ``ON_TOP``: the top cell, ``[cend - 1, cend)``;
- ``ON_INTERIOR_FACETS``: the facets between cells, ``[bottom, cend - 1)``
  (``[bottom, cend)`` on a periodic set); each map then names the entries
  of the cell below and of the cell above (twice the arity, below first).
"""
import enum

import numpy as np

__all__ = ["IterationRegion", "ON_BOTTOM", "ON_TOP", "ON_INTERIOR_FACETS", "ALL"]


class IterationRegion(enum.IntEnum):
    BOTTOM = 1
    TOP = 2
    INTERIOR_FACETS = 3
    ALL = 4


ON_BOTTOM = IterationRegion.BOTTOM
ON_TOP = IterationRegion.TOP
ON_INTERIOR_FACETS = IterationRegion.INTERIOR_FACETS
ALL = IterationRegion.ALL


def normalise(region):
    """``region`` as an IterationRegion (None is ALL)."""
    if region is None:
        return ALL
    try:
        return IterationRegion(region)
    except ValueError:
        raise ValueError("Unknown iteration region %r" % (region,)) from None


def facet_factor(region):
    """Cells a map entry spans: 2 for interior facets, else 1."""
    return 2 if normalise(region) == ON_INTERIOR_FACETS else 1


def layer_steps(ncell, region, periodic=False):
    """Per column with ``ncell`` cells, the first and one-past-last offset
    above the column's bottom that ``region`` runs (int64 arrays).

    A column with no cells runs nothing on top (its top cell would be the
    layer below its bottom); on the bottom it runs its bottom layer, as the
    reference does.  A periodic column has as many interior facets as cells
    (the last joins the top cell to the bottom one)."""
    ncell = np.asarray(ncell, dtype=np.int64)
    zero = np.zeros_like(ncell)
    region = normalise(region)
    if region == ALL:
        return zero, ncell
    if region == ON_BOTTOM:
        return zero, zero + 1
    if region == ON_TOP:
        return np.maximum(ncell - 1, 0), ncell
    if periodic:
        return zero, ncell
    return zero, np.maximum(ncell - 1, 0)
