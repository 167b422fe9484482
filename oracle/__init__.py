"""FP64 CPU oracle for the Nedelec assembly / CSR / solve path.

TEST INFRASTRUCTURE ONLY.  Nothing in the product package imports this
module.  Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline``
leg of ``bench.py`` may use it, and only as the checker (or as the timed CPU
baseline), never as the thing measured or shipped.

The oracle is a numpy/scipy restatement written from the mathematics of
first- and second-order Nedelec (first kind) elements and from the
conventions of the reference (DOLFIN 1.0 UnitCube, first-encounter entity
numbering, FIAT nodal dual bases).  Its pinning is recorded function by
function in ``oracle/fem.py``.
"""
from .fem import *  # noqa: F401,F403
