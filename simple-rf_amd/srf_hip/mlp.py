"""Fused PE + NeRF MLP on the HIP kernels, with a hand-written backward.

Mirrors class MLP (src/models/SimpleNeRF17.py:616-785): the same parameter names and shapes
(so state_dicts interchange), the same three variants (main, points-augmented with the 21/42
PE split, views-augmented with the 4-channel head).  The FP32 master parameters are repacked
into MFMA operand order on every call (srf_mlp_pack, ~1 MB), and the MLP runs with BF16,
FP16, FP32 or split-BF16 ('bf16_split': FP32 activations, products on BF16 MFMAs over operands
split into BF16 planes, three in the forward and two in the backward) MFMA operands and FP32
accumulation.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import torch

from . import _lib
from ._checks import dev_tensor, zero_grads

PREC = {'fp32': 0, 'float32': 0, 'bf16': 1, 'bfloat16': 1, 'fp16': 2, 'float16': 2, 'bf16_split': 3}

PARAM_NAMES = [f'pts_linears.{i}' for i in range(8)] + ['feature_linear', 'pts_output_linear', 'views_linears.0',
                                                        'views_output_linear']


class _Desc(C.Structure):
    _fields_ = [('trunk_pe', C.c_int), ('pts_pe_degree', C.c_int), ('dirs_pe_degree', C.c_int),
                ('view_dependent', C.c_int), ('sigma_softplus', C.c_int), ('prec', C.c_int)]


@dataclass(frozen=True)
class MLPSpec:
    """Structure of one MLP (reference config keys in comments)."""
    pts_pe_degree: int = 10          # points_positional_encoding_degree
    dirs_pe_degree: int = 4          # views_positional_encoding_degree
    sigma_pe_degree: int | None = None  # points_sigma_positional_encoding_degree (points-augmented)
    view_dependent: bool = True      # use_view_dirs and view_dependent_rgb
    sigma_softplus: bool = False
    prec: str = 'bf16'

    @property
    def trunk_pe(self) -> int:
        deg = self.pts_pe_degree if self.sigma_pe_degree is None else self.sigma_pe_degree
        return 3 * (2 * deg + 1)

    def desc(self) -> _Desc:
        if self.prec not in PREC:
            raise _lib.SrfError(f'unknown MLP precision {self.prec!r}')
        return _Desc(self.trunk_pe, self.pts_pe_degree, self.dirs_pe_degree, int(self.view_dependent),
                     int(self.sigma_softplus), PREC[self.prec])

    def present(self, i: int) -> bool:
        return self.view_dependent or i not in (8, 10, 11)

    def param_shapes(self, width=256, views_width=128):
        pts_dim = 3 * (2 * self.pts_pe_degree + 1)
        d = self.trunk_pe
        views_in = (pts_dim - d) + 3 * (2 * self.dirs_pe_degree + 1)
        shapes = []
        for i in range(8):
            k = d if i == 0 else (width + d if i == 5 else width)
            shapes.append(((width, k), (width,)))
        shapes.append(((width, width), (width,)))
        shapes.append(((1, width), (1,)) if self.view_dependent else ((4, width), (4,)))
        shapes.append(((views_width, width + views_in), (views_width,)))
        shapes.append(((3, views_width), (3,)))
        return [s if self.present(i) else None for i, s in enumerate(shapes)]


def _ptr_array(ts):
    arr = (C.c_void_p * len(ts))()
    for i, t in enumerate(ts):
        arr[i] = None if t is None else t.data_ptr()
    return arr


def packed_bytes(spec: MLPSpec) -> int:
    d = spec.desc()
    n = _lib.lib().srf_mlp_packed_bytes(C.byref(d))
    if n == 0:
        raise _lib.SrfError(f'unsupported MLP structure {spec}')
    return n


def workspace_bytes(spec: MLPSpec, n_points: int) -> int:
    d = spec.desc()
    return _lib.lib().srf_mlp_workspace_bytes(C.byref(d), int(n_points))


def pack(spec: MLPSpec, weights, biases, out: torch.Tensor | None = None) -> torch.Tensor:
    """Repack the 12 reference Linear weights/biases (None where absent) into MFMA order."""
    d = spec.desc()
    shapes = spec.param_shapes()
    ws, bs = [], []
    dev = None
    for i in range(12):
        if shapes[i] is None:
            ws.append(None)
            bs.append(None)
            continue
        w = dev_tensor(weights[i].detach(), PARAM_NAMES[i] + '.weight', shapes[i][0])
        b = dev_tensor(biases[i].detach(), PARAM_NAMES[i] + '.bias', shapes[i][1])
        ws.append(w)
        bs.append(b)
        dev = w.device
    nbytes = packed_bytes(spec)
    if out is None or out.numel() < nbytes:
        out = torch.empty(nbytes, dtype=torch.uint8, device=dev)
    _lib.call('srf_mlp_pack', C.byref(d), _ptr_array(ws), _ptr_array(bs), out.data_ptr(), _lib.stream_ptr(dev))
    return out


class _MLPFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, spec, train, pts, dirs, noise, rays, *params):
        # pts [P,3] (dirs per point), or pts None and rays = (rays_o, rays_d, z_vals) with the points
        # generated in the kernel (dirs per ray)
        weights, biases = params[:12], params[12:]
        packed = pack(spec, weights, biases)
        if rays is None:
            n, dev = pts.shape[0], pts.device
        else:
            n, dev = rays[2].numel(), rays[2].device
        sigma = torch.empty((n,), dtype=torch.float32, device=dev)
        rgb = torch.empty((n, 3), dtype=torch.float32, device=dev)
        ws = torch.empty(workspace_bytes(spec, n), dtype=torch.uint8, device=dev) if train else None
        d = spec.desc()
        if rays is None:
            _lib.call('srf_mlp_forward', C.byref(d), packed.data_ptr(), n, pts.data_ptr(), _lib.ptr(dirs),
                      _lib.ptr(noise), sigma.data_ptr(), rgb.data_ptr(), _lib.ptr(ws), _lib.stream_ptr(dev))
        else:
            o, dr, z = rays
            _lib.call('srf_mlp_forward_rays', C.byref(d), packed.data_ptr(), z.shape[0], z.shape[1], o.data_ptr(),
                      dr.data_ptr(), z.data_ptr(), _lib.ptr(dirs), _lib.ptr(noise), sigma.data_ptr(),
                      rgb.data_ptr(), _lib.ptr(ws), _lib.stream_ptr(dev))
        ctx.spec = spec
        ctx.shapes = [None if p is None else p.shape for p in params]
        ctx.save_for_backward(packed, ws, sigma, rgb)
        return sigma, rgb

    @staticmethod
    def backward(ctx, d_sigma, d_rgb):
        packed, ws, sigma, rgb = ctx.saved_tensors
        spec = ctx.spec
        if ws is None:
            raise _lib.SrfError('MLP backward without a training workspace')
        dev = sigma.device
        n = sigma.shape[0]
        d_sigma = torch.zeros_like(sigma) if d_sigma is None else d_sigma.contiguous()
        d_rgb = torch.zeros_like(rgb) if d_rgb is None else d_rgb.contiguous()
        grads = zero_grads(ctx.shapes, dev)  # srf_mlp_weight_grads accumulates
        d = spec.desc()
        stream = _lib.stream_ptr(dev)
        _lib.call('srf_mlp_backward_data', C.byref(d), packed.data_ptr(), n, sigma.data_ptr(), rgb.data_ptr(),
                  d_sigma.data_ptr(), d_rgb.data_ptr(), ws.data_ptr(), stream)
        _lib.call('srf_mlp_weight_grads', C.byref(d), n, ws.data_ptr(), _ptr_array(grads[:12]),
                  _ptr_array(grads[12:]), stream)
        return (None, None, None, None, None, None, *grads)


def mlp_forward(spec: MLPSpec, weights, biases, pts: torch.Tensor, dirs: torch.Tensor | None,
                noise: torch.Tensor | None = None):
    """sigma [P], rgb [P,3] of the MLP at points pts [P,3] (and view dirs [P,3]).

    ``weights``/``biases``: 12-entry lists in PARAM_NAMES order (None where the variant has
    no such layer).  ``noise`` [P] is the already-scaled sigma noise or None.
    Differentiable w.r.t. the parameters (pts/dirs/noise get no gradient, as in the reference
    where points are not optimised)."""
    pts = dev_tensor(pts, 'pts')
    n = pts.shape[0]
    if pts.shape != (n, 3):
        raise _lib.SrfError(f'pts must be [P,3], got {tuple(pts.shape)}')
    if spec.view_dependent:
        dirs = dev_tensor(dirs, 'dirs', (n, 3))
    else:
        dirs = None
    if noise is not None:
        noise = dev_tensor(noise.reshape(-1), 'noise', (n,))
    params = [None if spec.param_shapes()[i] is None else weights[i] for i in range(12)] + \
             [None if spec.param_shapes()[i] is None else biases[i] for i in range(12)]
    train = torch.is_grad_enabled() and any(p is not None and p.requires_grad for p in params)
    return _MLPFunction.apply(spec, train, pts, dirs, noise, None, *params)


def mlp_forward_rays(spec: MLPSpec, weights, biases, rays_o: torch.Tensor, rays_d: torch.Tensor,
                     z_vals: torch.Tensor, view_dirs: torch.Tensor | None, noise: torch.Tensor | None = None):
    """mlp_forward at the points rays_o[:, None] + rays_d[:, None] * z_vals[..., None], generated in the
    kernel (src/models/SimpleNeRF17.py:213-215, :266-268) instead of materialised: the same sigma
    [R*S] and rgb [R*S,3] as mlp_forward on those points, bit for bit.  view_dirs [R,3] is one
    direction per ray (run_network's expand, :429-433).  The rays get no gradient, as pts in
    mlp_forward."""
    z_vals = dev_tensor(z_vals, 'z_vals')
    if z_vals.dim() != 2:
        raise _lib.SrfError(f'z_vals must be [R,S], got {tuple(z_vals.shape)}')
    r, s = z_vals.shape
    rays_o = dev_tensor(rays_o, 'rays_o', (r, 3))
    rays_d = dev_tensor(rays_d, 'rays_d', (r, 3))
    if spec.view_dependent:
        view_dirs = dev_tensor(view_dirs, 'view_dirs', (r, 3))
    else:
        view_dirs = None
    if noise is not None:
        noise = dev_tensor(noise.reshape(-1), 'noise', (r * s,))
    params = [None if spec.param_shapes()[i] is None else weights[i] for i in range(12)] + \
             [None if spec.param_shapes()[i] is None else biases[i] for i in range(12)]
    train = torch.is_grad_enabled() and any(p is not None and p.requires_grad for p in params)
    return _MLPFunction.apply(spec, train, None, view_dirs, noise, (rays_o, rays_d, z_vals), *params)
