"""FP64 CPU oracle of kUPS's batched particle-interaction hot path.

TEST INFRASTRUCTURE ONLY.  This package is the parity checker of the HIP
library and the CPU baseline timed next to it.  Only ``tests/``,
``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg import it;
the product package ``kups_amd`` never does.

The compiled part (``kups_oracle.c`` -> ``liboracle.so``) restates the
reference's algorithms in FP64; ``oracle.py`` holds the numpy restatements
(Ewald parameter estimation, GCMC log-ratios, Peng-Robinson fugacity, the
LJ tail correction) and the ctypes bindings.  Each function cites the
reference file:line it follows.  Parity pins for every function are listed
in ``tests/test_oracle.py``.
"""

from .oracle import *  # noqa: F401,F403
