"""Operator generator: index layout, straight-line operator programs,
host C/C++ and HIP printers, and the ``generate_code`` entry point."""
from .layout import Nterms, Msize, Lsize, generate_mappings, idx, monomials
from .writer import generate_code

__all__ = ["Nterms", "Msize", "Lsize", "generate_mappings", "idx", "monomials", "generate_code"]
