"""Fermion operators of gpt_amd (lib/gpt/qcd/fermion/__init__.py)."""
from .operator import wilson_clover, fine_operator, apply_open_boundaries  # noqa: F401
from . import preconditioner  # noqa: F401
