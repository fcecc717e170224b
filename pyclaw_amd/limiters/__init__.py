"""Limiter ids (src/pyclaw/limiters/tvd.py:74-77)."""
from . import tvd

__all__ = ['tvd']
