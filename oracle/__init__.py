"""The oracle: CPU checker of the HIP backend (test infrastructure only).

``Oracle(prec)`` binds the C restatement (liboracle_{f32,f64}.so).
``Reference(prec)`` binds the reference's own sources compiled here
(oracle/_ref/libfca_ref_{f32,f64}.so), present only where the reference tree
was available at build time. Only tests/, __graft_entry__.smoke and bench.py's
CPU baseline import this package.
"""
from .binding import Oracle, Reference, ref_available, ORACLE_DIR  # noqa: F401
