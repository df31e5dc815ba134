"""CPU oracle for line2Dup's matching path -- test infrastructure only.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import this package, and only as the checker (or as the timed CPU baseline).  The product
library never calls it.
"""
from .py_oracle import *  # noqa: F401,F403
