"""FP64 NumPy restatement of the reference's EFF/OLA operator and the
algorithms built on it.

This package is TEST INFRASTRUCTURE.  Only ``tests/``, ``__graft_entry__.smoke()``
and the ``cpu_baseline`` leg of ``bench.py`` may import it, and only as the
checker (or as the timed CPU baseline).  The product path (``vmbd_amd``) never
imports, calls or falls back to anything in here.
"""
