"""FP64 CPU oracle of kUPS's batched particle-interaction hot path.

TEST INFRASTRUCTURE ONLY.  This package is the parity checker of the HIP
library and the CPU baseline timed next to it.  Only ``tests/``,
``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg import it;
the product package ``kups_amd`` never does.

The compiled part (``kups_oracle.c`` -> ``liboracle.so``) restates the
reference's algorithms in FP64: neighbor lists and rh-subset queries, LJ
(plain and switched), Morse, Ewald real and reciprocal space, the MD steps,
CSVR and stochastic cell rescaling, GCMC trial moves, Philox4x32-10.
``oracle.py`` holds the ctypes bindings and the numpy pieces of the MLP
reference (forward pass, loss and gradient).  Each function cites the
reference file:line it follows.  Parity pins for every function are in
``tests/test_oracle.py``, ``tests/test_gcmc_cpu.py``,
``tests/test_nl_query_cpu.py`` and ``tests/test_rng_barostat_cpu.py``.
"""

from .oracle import *  # noqa: F401,F403
