"""Fusion back end: super-op definitions -> HIP fused-operation libraries."""
from .codegen import build, generate  # noqa: F401
from .superops import parse  # noqa: F401
