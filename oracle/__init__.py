"""CPU oracle for the VEGAS hot path: test infrastructure only.

This package restates, in NumPy FP64 (and float32 where the reference does),
the reference algorithm of vegasflow 1.4.0 for the sample -> map -> integrand
-> digest -> refine iteration, so that the HIP kernels of ``vegasflow_amd``
can be checked on identical inputs.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import it, and only as the parity reference or as the
timed CPU baseline. The product path (``vegasflow_amd``) never imports it and has no
CPU fallback.
"""
