"""MI355X-native aggregation-AMG preconditioned Krylov solver (ThermoFEM's hot path).

The compute path is the in-tree HIP library libthermofem_hip.so (gfx950);
this package is its Python binding, the problem generators and the
reference-shaped solver interface.
"""
from ._lib import LIB_PATH, TfError, lib  # noqa: F401
