"""Test infrastructure only: the FP64 NumPy restatement of diffractom's single-phase forward operator and its
SPDHG-Poisson solver (``oracle.model``), and a C/OpenMP FP32 restatement used as the host-core baseline
(``oracle/c``). Only the tests, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may import
or run anything here, always as the checker; the product package ``diffractom_amd`` never does.
"""
