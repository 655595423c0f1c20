"""CPU FP64 oracle for the phase-field time step.

This package is test infrastructure: only ``tests/``, ``__graft_entry__.smoke()``
and ``bench.py``'s CPU-baseline leg import it, and only as the checker.  The
product library never calls into it.
"""
