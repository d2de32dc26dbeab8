"""CPU FP64 oracle for the MAM4 microphysics chain -- test infrastructure.

Only the tests, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` use this package, and only as the checker or the timed CPU
baseline.  The shipped library never calls it.
"""
from .pyoracle import Oracle, build, lib_path  # noqa: F401
