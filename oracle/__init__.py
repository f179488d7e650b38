"""CPU oracle for the ARTED real-time hot path -- TEST INFRASTRUCTURE ONLY.

* ``oracle.cpu``  -- ctypes wrapper of ``salmon_oracle.c``, a C restatement of
  the reference routines (file:line citations in that file).
* ``oracle.ref``  -- ctypes wrapper of ``_ref/libsalmon_ref.so``, the
  reference's own Fortran for this path compiled from its sources by
  ``oracle/ref/Makefile`` (built only where the reference tree is present).

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may use this package, and only as the checker or the timed CPU
baseline -- never as the product path.
"""
