"""CPU oracle for the Metropolis + parallel-tempering engine (test infrastructure).

This package is the checker: only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` import it.  The product path
(``gpu-monte-carlo_amd/``) never does.  The C restatement lives in
``mc_oracle.c``/``oracle_models.h``; :mod:`oracle.api` wraps it with numpy.
"""

from .api import *  # noqa: F401,F403
