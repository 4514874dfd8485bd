"""CPU oracle for the LU solve path -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import this package, and only as the checker. The product library never imports it.
"""
from .oracle import *  # noqa: F401,F403
