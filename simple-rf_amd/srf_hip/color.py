"""TensoRF colour network on the HIP kernels (basis matrix + MlpFeaturesColorPredictor), with the
hand-written backward.  Mirrors src/models/SimpleTensoRF09.py:1262-1270 and :1370-1421.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import torch

from . import _lib
from ._checks import dev_tensor, zero_grads


class _Desc(C.Structure):
    _fields_ = [('n_features', C.c_int), ('basis_dim', C.c_int), ('features_pe', C.c_int),
                ('use_view_dirs', C.c_int), ('views_pe', C.c_int), ('hidden', C.c_int), ('prec', C.c_int)]


# operand precision of the colour network's products (activations, storage, accumulation and the
# weight gradients stay FP32): FP32, split BF16 (three planes forward, two backward), plain BF16
PRECISIONS = {'fp32': 0, 'bf16': 1, 'bf16_split': 3}   # SRF_PREC_F32 / _BF16 / _BF16_SPLIT


@dataclass(frozen=True)
class ColorSpec:
    n_features: int          # sum of the appearance components
    basis_dim: int = 27      # features_dimension_color
    features_pe: int = 0     # features_positional_encoding_degree
    use_view_dirs: bool = True
    views_pe: int = 0        # views_positional_encoding_degree
    hidden: int = 128        # num_units_color_predictor
    prec: str = 'fp32'       # a key of PRECISIONS

    def desc(self) -> _Desc:
        if self.prec not in PRECISIONS:
            raise _lib.SrfError(f'colour precision {self.prec!r} is not one of {sorted(PRECISIONS)}')
        return _Desc(self.n_features, self.basis_dim, self.features_pe, int(self.use_view_dirs), self.views_pe,
                     self.hidden, PRECISIONS[self.prec])

    def supported(self) -> bool:
        d = self.desc()
        return _lib.lib().srf_color_packed_bytes(C.byref(d)) > 0


def pack(spec: ColorSpec, basis, w1, b1, w2, b2, w3, b3) -> torch.Tensor:
    d = spec.desc()
    nbytes = _lib.lib().srf_color_packed_bytes(C.byref(d))
    if nbytes == 0:
        raise _lib.SrfError(f'colour network {spec} is outside what the kernels support')
    ts = [dev_tensor(t.detach(), n) for t, n in zip((basis, w1, b1, w2, b2, w3, b3),
                                                    ('basis', 'w1', 'b1', 'w2', 'b2', 'w3', 'b3'))]
    out = torch.empty((nbytes // 4,), dtype=torch.float32, device=basis.device)
    _lib.call('srf_color_pack', C.byref(d), *[_lib.ptr(t) for t in ts], _lib.ptr(out), _lib.stream_ptr(basis.device))
    return out


class _ColorFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, spec, train, app, idx, count, dense_rows, n_samples, view_dirs, basis, w1, b1, w2, b2, w3, b3):
        d = spec.desc()
        n = app.shape[0]
        dev = app.device
        packed = pack(spec, basis, w1, b1, w2, b2, w3, b3)
        ws = None
        if train:
            ws = torch.empty((_lib.lib().srf_color_workspace_bytes(C.byref(d), n),), dtype=torch.uint8, device=dev)
        at_idx = dense_rows is not None
        if at_idx:
            rgb = torch.zeros((dense_rows, 3), dtype=torch.float32, device=dev)
        else:
            rgb = torch.empty((n, 3), dtype=torch.float32, device=dev)
        _lib.call('srf_color_forward', C.byref(d), _lib.ptr(packed), n, _lib.ptr(count), _lib.ptr(app), _lib.ptr(idx),
                  int(n_samples), _lib.ptr(view_dirs), _lib.ptr(rgb), int(at_idx), _lib.ptr(ws), _lib.stream_ptr(dev))
        ctx.spec = spec
        ctx.n = n
        ctx.at_idx = at_idx
        ctx.save_for_backward(packed, ws, rgb, idx, count, basis, w1, b1, w2, b2, w3, b3)
        return rgb

    @staticmethod
    def backward(ctx, g_rgb):
        packed, ws, rgb, idx, count, basis, w1, b1, w2, b2, w3, b3 = ctx.saved_tensors
        if ws is None:
            raise _lib.SrfError('colour backward needs a training forward (workspace)')
        spec = ctx.spec
        d = spec.desc()
        n = ctx.n
        d_app = torch.empty((n, spec.n_features), dtype=torch.float32, device=rgb.device)
        grads = zero_grads([t.shape for t in (basis, w1, b1, w2, b2, w3, b3)], basis.device)
        _lib.call('srf_color_backward', C.byref(d), _lib.ptr(packed), n, _lib.ptr(count), _lib.ptr(idx),
                  int(ctx.at_idx), _lib.ptr(rgb), _lib.ptr(g_rgb.contiguous()), _lib.ptr(ws), _lib.ptr(d_app),
                  *[_lib.ptr(g) for g in grads], _lib.stream_ptr(rgb.device))
        return (None, None, d_app, None, None, None, None, None, *grads)


def color_forward(spec: ColorSpec, app: torch.Tensor, basis, mlp_params, *, view_dirs: torch.Tensor | None = None,
                  idx: torch.Tensor | None = None, n_samples: int = 1, count: torch.Tensor | None = None,
                  dense_rows: int | None = None) -> torch.Tensor:
    """rgb [M,3] = sigmoid(MLP([PE(basis app), PE(dir)])) for appearance rows app [M, n_features].

    ``mlp_params``: (w1, b1, w2, b2, w3, b3) of color_predictor.mlp.{0,2,4}.  The view direction of
    row j is view_dirs[(idx[j] if idx is not None else j) // n_samples] (view_dirs [rays, 3]).
    ``count`` (device int32 [1]): only the rows below it are computed (M is then a capacity), with no
    host synchronisation.  ``dense_rows``: return the dense sample buffer [dense_rows, 3] instead,
    row idx[j] holding row j's colour and the other rows 0 (the reference's rgb = zeros(...),
    rgb[validity_mask] = surface_rgb, SimpleTensoRF09.py:1246, :1271)."""
    app = dev_tensor(app, 'app')
    if app.dim() != 2 or app.shape[1] != spec.n_features:
        raise _lib.SrfError(f'app must be [M, {spec.n_features}], got {tuple(app.shape)}')
    if app.data_ptr() % 16:
        app = app.clone()
    if idx is not None:
        idx = dev_tensor(idx, 'idx', dtype=torch.int32)
    if count is not None:
        count = dev_tensor(count, 'count', (1,), dtype=torch.int32)
    if dense_rows is not None and idx is None:
        raise _lib.SrfError('dense_rows needs idx')
    if spec.use_view_dirs:
        view_dirs = dev_tensor(view_dirs, 'view_dirs')
    else:
        view_dirs = None
    w1, b1, w2, b2, w3, b3 = mlp_params
    params = (basis, w1, b1, w2, b2, w3, b3)
    train = torch.is_grad_enabled() and (app.requires_grad or any(p.requires_grad for p in params))
    return _ColorFunction.apply(spec, train, app, idx, count, None if dense_rows is None else int(dense_rows),
                                int(n_samples), view_dirs, *params)
