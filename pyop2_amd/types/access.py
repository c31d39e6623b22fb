"""Access descriptors of par_loop arguments.

The six modes and their integer values are the reference's interface
(``pyop2/types/access.py:4-37``), kept so that user code and the argument
signatures of cached loops are unchanged.  What each mode means for the
device loops:

* ``READ``  -- the kernel only gathers the values;
* ``WRITE`` -- the kernel's packed values replace the target entries (an
  indirect WRITE whose targets collide is coloured, so the last element in
  loop order wins as in a sequential loop);
* ``RW``    -- gathered, then written back (coloured when indirect);
* ``INC``   -- the kernel's packed values start at zero and are added to the
  target (Dat, Global or Mat); the device resolves the conflicts;
* ``MIN`` / ``MAX`` -- reductions into a Global over all elements.
"""
import enum


class Access(enum.IntEnum):
    READ = 1
    WRITE = 2
    RW = 3
    INC = 4
    MIN = 5
    MAX = 6


READ = Access.READ
WRITE = Access.WRITE
RW = Access.RW
INC = Access.INC
MIN = Access.MIN
MAX = Access.MAX
