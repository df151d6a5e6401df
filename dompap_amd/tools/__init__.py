"""Tools mirroring dompap.tools."""
from .autotune import autotune  # noqa: F401
