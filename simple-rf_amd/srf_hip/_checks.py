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


def zero_grads(shapes, device) -> list:
    """Zero-filled FP32 gradient buffers of the given shapes (None entries stay None), as views of
    one buffer: a single fill launch instead of one per parameter.  Each view starts on a 64-byte
    boundary (vectorised optimizer kernels keep their aligned path)."""
    sizes = [0 if s is None else int(torch.Size(s).numel()) for s in shapes]
    padded = [(n + 15) // 16 * 16 for n in sizes]
    flat = torch.zeros(max(sum(padded), 1), dtype=torch.float32, device=device)
    out, o = [], 0
    for s, n, pn in zip(shapes, sizes, padded):
        out.append(None if s is None else flat[o:o + n].view(s))
        o += pn
    return out
