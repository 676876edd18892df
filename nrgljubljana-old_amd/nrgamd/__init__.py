"""Python side of the MI355X NRG engine: table derivation, `data` generation,
and the ctypes binding of the C-ABI library (used by tests and bench.py)."""
from . import symtab, datagen  # noqa: F401

__all__ = ["symtab", "datagen", "capi"]
