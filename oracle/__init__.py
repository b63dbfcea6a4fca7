"""CPU oracle for the freud pair-analysis path -- TEST INFRASTRUCTURE ONLY.

A C++ restatement of freud 3.5.0's arithmetic (``freud_oracle.cpp``) and a
ctypes wrapper around it.  Only ``tests/``, ``__graft_entry__.smoke()`` and
``bench.py``'s ``cpu_baseline`` leg may import this package; the product path
(``freud_amd``) never does.
"""

from .api import *  # noqa: F401,F403
