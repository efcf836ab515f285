"""MI355X (gfx950) HIP implementation of Simple-RF's rendering hot path.

Host side of the drop-in: torch-facing wrappers (with hand-written backwards) around the
C-ABI library libsrf_hip.so built from ../csrc.  There is a single code path: every op
runs the HIP kernels and raises if the library is unavailable.
"""
from ._lib import SrfError, lib  # noqa: F401

__all__ = ['SrfError', 'lib']
