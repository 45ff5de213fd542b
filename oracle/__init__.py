"""CPU oracle for the TBPM hot path -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this package, and only as the checker.  The product
(``tbplas_amd``) never imports it and has no CPU fallback.

Contents (each function cites the reference file:line it restates):

* ``ref_*``      -- the reference's own Fortran TBPM subroutines, compiled from
                    ``tbplas/fortran/*.f90`` by ``oracle/Makefile`` into
                    ``oracle/_ref/libtbpm_ref.so`` and called through their
                    Fortran symbols (no f2py wrapper).
* ``orc_*``      -- the C restatement in ``oracle/tbpm_oracle.c``
                    (``oracle/_build/liboracle.so``), fed caller-supplied
                    initial states.
* numpy/Python restatements of the host-side pieces: CSR builders, Bessel
  series, graphene supercell hop lists, Peierls phases, rescale estimate and
  the Analyzer formulas.
* independent exact oracles: k-space propagation of pristine periodic
  graphene and dense eigendecomposition for small samples.

How each function is pinned (closed forms, golden vectors, the compiled
reference) is listed next to it; nothing here is "parity unpinned" except
where a docstring says so.
"""
from .core import *  # noqa: F401,F403
