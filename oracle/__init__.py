"""CPU oracle for the P3M Yukawa force path.

This package is test infrastructure: a plain restatement of the reference
algorithm used only to check the HIP path.  Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import it.  The product path (``mdforce``) never imports it and has no CPU
fallback.
"""
