"""FP64 CPU oracle for the PO far-field path (test infrastructure only).

Only tests/, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` import this package, and only as the checker.  The product
library (``wavefiber3d_amd``) never imports it.
"""
from .oracle import *  # noqa: F401,F403
