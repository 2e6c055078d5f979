"""CPU oracles for the amigo_amd hot path (test infrastructure only).

- ``oracle.ref``: the reference's own SparseLDL/SparseCholesky/BlockAMD
  compiled from its headers (``make -C oracle``) into ``oracle/_ref``.
- ``oracle.ldl``: this repository's C restatement of the reference's numeric
  multifrontal LDL^T (``oracle_ldl.c``).
- ``oracle.ipm``: numpy restatement of the reference's ``ipm::`` kernels
  (include/interior_point_backend.h) and of the optimizer wrappers.

Only tests/, ``__graft_entry__.smoke()`` and bench.py's cpu_baseline leg may
import this package, and only as the checker.
"""
