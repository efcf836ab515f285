"""Argument validation shared by the torch-facing wrappers."""
from __future__ import annotations

import torch

from ._lib import SrfError


def dev_tensor(t: torch.Tensor, name: str, shape=None, dtype=torch.float32) -> torch.Tensor:
    """Require a contiguous tensor of ``dtype`` on a GPU device (and ``shape`` if given)."""
    if not isinstance(t, torch.Tensor):
        raise SrfError(f'{name}: expected a torch.Tensor, got {type(t).__name__}')
    if t.device.type != 'cuda':
        raise SrfError(f'{name}: expected a GPU tensor, got device {t.device} '
                       '(the HIP kernels are the only implementation)')
    if t.dtype != dtype:
        raise SrfError(f'{name}: expected {dtype}, got {t.dtype}')
    if shape is not None and tuple(t.shape) != tuple(shape):
        raise SrfError(f'{name}: expected shape {tuple(shape)}, got {tuple(t.shape)}')
    return t.contiguous()


def opt_tensor(t, name, shape=None, dtype=torch.float32):
    return None if t is None else dev_tensor(t, name, shape, dtype)
