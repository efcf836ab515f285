"""TensoRF factor sampling (vector-matrix and CP decompositions) on the HIP kernels, with
hand-written backwards.

Mirrors LowRankTensor.forward's validity and normalisation, AlphaGridMask.sample_alpha,
VmDecomposedTensor.get_volume_density / get_color, CpDecomposedTensor.get_volume_density /
get_color (src/models/SimpleTensoRF09.py:701-765, 1214-1263, 1043-1089, 1342-1345), the surface
test of :726 and the non-NDC sampler of :382-400.

Factor layout: the kernels read plane p as [H, W, c] and line p as [L, c] (component axis
innermost).  Parameters keep the reference's shapes ([1, c, H, W] and [1, c, L, 1], so
state_dict keys and shapes are unchanged) but are created as permuted views of texel-major
storage (:func:`texel_major_param`); a factor in the reference's own contiguous layout is
accepted too and transposed on the fly.  Gradients come back in the parameter's layout.
"""
from __future__ import annotations

import ctypes as C

import torch

from . import _lib
from ._checks import dev_tensor

RELU = 0
SOFTPLUS = 1
_PREDICTORS = {'ReLU': RELU, 'SoftPlus': SOFTPLUS}


def predictor_code(name: str) -> int:
    """Density predictor name of the tensor configs (SimpleTensoRF09.py:671-679) -> ABI code."""
    if name not in _PREDICTORS:
        raise NotImplementedError(name)
    return _PREDICTORS[name]


def texel_major_param(n_comp: int, rows: int, cols: int, init: torch.Tensor | None = None,
                      device=None) -> torch.nn.Parameter:
    """A [1, n_comp, rows, cols] parameter whose storage is texel-major [rows, cols, n_comp]."""
    store = torch.empty((1, rows, cols, n_comp), dtype=torch.float32, device=device)
    t = store.permute(0, 3, 1, 2)
    if init is not None:
        with torch.no_grad():
            t.copy_(init)
    return torch.nn.Parameter(t)


def _texel_view(t: torch.Tensor, name: str) -> torch.Tensor:
    """[1, c, H, W] factor -> contiguous [H, W, c] view (a copy only if t is not texel-major)."""
    if t.dim() != 4 or t.shape[0] != 1:
        raise _lib.SrfError(f'{name}: expected a [1, c, H, W] factor, got {tuple(t.shape)}')
    if t.device.type != 'cuda' or t.dtype != torch.float32:
        raise _lib.SrfError(f'{name}: expected a float32 GPU tensor (the HIP kernels are the only implementation)')
    v = t[0].permute(1, 2, 0)
    if not v.is_contiguous():
        v = v.contiguous()
    if v.data_ptr() % 16:
        v = v.clone()
    return v


def _factors_struct(planes, lines):
    """struct srf_vm_factors of three planes and three lines (plus the views it points into)."""
    f = _lib.VmFactors()
    keep = []
    for p in range(3):
        P = _texel_view(planes[p], f'plane {p}')
        L = _texel_view(lines[p], f'line {p}')
        c = P.shape[2]
        if L.shape[2] != c or L.shape[1] != 1:
            raise _lib.SrfError(f'line {p}: expected [1, {c}, L, 1], got {tuple(lines[p].shape)}')
        f.n_comp[p] = c
        f.plane_h[p] = P.shape[0]
        f.plane_w[p] = P.shape[1]
        f.line_len[p] = L.shape[0]
        f.planes[p] = P.data_ptr()
        f.lines[p] = L.data_ptr()
        keep += [P, L]
    return f, keep


def _grad_buffers(planes, lines, needs):
    """Zeroed texel-major gradient buffers for the factors whose grad is needed (others get a
    scratch buffer the kernels may write into)."""
    dp, dl, outs = [], [], []
    for p in range(3):
        for src, dst in ((planes[p], dp), (lines[p], dl)):
            _, c, h, w = src.shape
            g = torch.zeros((1, h, w, c), dtype=torch.float32, device=src.device)
            dst.append(g)
    arr_p = (C.c_void_p * 3)(*[g.data_ptr() for g in dp])
    arr_l = (C.c_void_p * 3)(*[g.data_ptr() for g in dl])
    grads = []
    for p in range(3):
        grads.append(dp[p].permute(0, 3, 1, 2) if needs[p] else None)
    for p in range(3):
        grads.append(dl[p].permute(0, 3, 1, 2) if needs[3 + p] else None)
    return arr_p, arr_l, grads


def _bbox_arg(bbox) -> C.Array:
    vals = [float(x) for x in bbox]
    if len(vals) != 6:
        raise _lib.SrfError(f'bbox: expected 6 floats (lo xyz, hi xyz), got {len(vals)}')
    return (C.c_float * 6)(*vals)


class AlphaMask:
    """AlphaGridMask's occupancy volume for the kernels (SimpleTensoRF09.py:1323-1349):
    volume [D, H, W] float on the GPU and its bounding box (6 host floats)."""

    def __init__(self, volume: torch.Tensor, bbox):
        vol = volume.reshape(volume.shape[-3:]).to(torch.float32)
        self.volume = dev_tensor(vol, 'alpha volume')
        self.bbox = [float(x) for x in bbox]

    def struct(self):
        g = _lib.AlphaGrid()
        g.volume = self.volume.data_ptr()
        g.depth, g.height, g.width = self.volume.shape
        for k in range(6):
            g.bbox[k] = self.bbox[k]
        return g


def alpha_mask_sample(mask: AlphaMask, pts: torch.Tensor) -> torch.Tensor:
    """AlphaGridMask.sample_alpha: trilinear occupancy of pts [M,3] -> [M]."""
    pts = dev_tensor(pts, 'pts').reshape(-1, 3)
    out = torch.empty((pts.shape[0],), dtype=torch.float32, device=pts.device)
    g = mask.struct()
    _lib.call('srf_alpha_mask_sample', C.byref(g), pts.shape[0], _lib.ptr(pts), _lib.ptr(out),
              _lib.stream_ptr(pts.device))
    return out


# The VM factor backwards order the contributing samples by plane cell before the scatter (a
# workspace from srf_vm_bwd_workspace_bytes); False walks them in sample order instead (same sums up
# to the order of the FP32 atomics, more atomics).
SORTED_BACKWARD = True


def _sorted_ws(f, rows: int, device):
    if not SORTED_BACKWARD:
        return None
    nbytes = _lib.lib().srf_vm_bwd_workspace_bytes(C.byref(f), int(rows))
    return torch.empty((max(int(nbytes), 1),), dtype=torch.uint8, device=device)


class _VmDensity(torch.autograd.Function):
    @staticmethod
    def forward(ctx, pts, bbox, alpha, predictor, offset, *factors):
        planes, lines = factors[:3], factors[3:]
        f, keep = _factors_struct(planes, lines)
        n = pts.shape[0]
        dev = pts.device
        sigma = torch.empty((n,), dtype=torch.float32, device=dev)
        feat = torch.empty((n,), dtype=torch.float32, device=dev)
        valid = torch.empty((n,), dtype=torch.uint8, device=dev)
        g = alpha.struct() if alpha is not None else None
        _lib.call('srf_vm_density_fwd', C.byref(f), _bbox_arg(bbox), C.byref(g) if g is not None else None, n,
                  _lib.ptr(pts), predictor, float(offset), _lib.ptr(sigma), _lib.ptr(feat), _lib.ptr(valid),
                  _lib.stream_ptr(dev))
        ctx.save_for_backward(pts, feat, valid, *factors)
        ctx.cfg = (bbox, predictor, float(offset))
        ctx.mark_non_differentiable(valid)
        return sigma, valid

    @staticmethod
    def backward(ctx, g_sigma, _g_valid):
        pts, feat, valid, *factors = ctx.saved_tensors
        bbox, predictor, offset = ctx.cfg
        needs = ctx.needs_input_grad[5:]
        if not any(needs) or g_sigma is None:
            return (None,) * 11
        planes, lines = factors[:3], factors[3:]
        f, keep = _factors_struct(planes, lines)
        dp, dl, grads = _grad_buffers(planes, lines, needs)
        ws = _sorted_ws(f, pts.shape[0], pts.device)
        _lib.call('srf_vm_density_bwd', C.byref(f), dp, dl, _bbox_arg(bbox), pts.shape[0], _lib.ptr(pts),
                  _lib.ptr(valid), _lib.ptr(feat), predictor, offset, _lib.ptr(g_sigma.contiguous()), _lib.ptr(ws),
                  _lib.stream_ptr(pts.device))
        return (None, None, None, None, None, *grads)


def vm_density(pts: torch.Tensor, planes, lines, bbox, *, alpha: AlphaMask | None = None,
               predictor: int = RELU, offset: float = 0.0):
    """Density of every sample: sigma [...] and validity [...] (uint8) for pts [..., 3].

    Invalid samples (outside bbox or, with ``alpha``, unoccupied) get sigma 0, as
    LowRankTensor.forward + get_volume_density leave them (SimpleTensoRF09.py:701-721,
    :1214-1239).  Gradients flow to the six factors; the coordinates are detached.
    """
    shape = pts.shape[:-1]
    p = dev_tensor(pts.detach(), 'pts').reshape(-1, 3)
    sigma, valid = _VmDensity.apply(p, tuple(float(x) for x in bbox), alpha, int(predictor), float(offset),
                                    *planes, *lines)
    return sigma.view(shape), valid.view(shape)


class _VmAppearance(torch.autograd.Function):
    @staticmethod
    def forward(ctx, pts, idx, n_rows, count, bbox, *factors):
        planes, lines = factors[:3], factors[3:]
        f, keep = _factors_struct(planes, lines)
        total = sum(f.n_comp)
        out = torch.empty((n_rows, total), dtype=torch.float32, device=pts.device)
        _lib.call('srf_vm_appearance_fwd', C.byref(f), _bbox_arg(bbox), n_rows, _lib.ptr(count), _lib.ptr(idx),
                  _lib.ptr(pts), _lib.ptr(out), _lib.stream_ptr(pts.device))
        ctx.save_for_backward(pts, idx, count, *factors)
        ctx.cfg = (n_rows, bbox)
        return out

    @staticmethod
    def backward(ctx, g_out):
        pts, idx, count, *factors = ctx.saved_tensors
        n_rows, bbox = ctx.cfg
        needs = ctx.needs_input_grad[5:]
        if not any(needs):
            return (None,) * 11
        planes, lines = factors[:3], factors[3:]
        f, keep = _factors_struct(planes, lines)
        dp, dl, grads = _grad_buffers(planes, lines, needs)
        g = g_out.contiguous()
        if g.data_ptr() % 16:   # rows are read as float4
            g = g.clone()
        ws = _sorted_ws(f, n_rows, pts.device)
        _lib.call('srf_vm_appearance_bwd', C.byref(f), dp, dl, _bbox_arg(bbox), n_rows, _lib.ptr(count),
                  _lib.ptr(idx), _lib.ptr(pts), _lib.ptr(g), _lib.ptr(ws), _lib.stream_ptr(pts.device))
        return (None, None, None, None, None, *grads)


def _rows_and_count(p, idx, n_rows, count):
    if idx is None:
        n_rows = p.shape[0]
    else:
        idx = dev_tensor(idx, 'idx', dtype=torch.int32)
        n_rows = idx.shape[0] if n_rows is None else int(n_rows)
    if count is not None:
        count = dev_tensor(count, 'count', (1,), dtype=torch.int32)
    return idx, n_rows, count


def vm_appearance(pts: torch.Tensor, planes, lines, bbox, idx: torch.Tensor | None = None, n_rows: int | None = None,
                  count: torch.Tensor | None = None):
    """Appearance features cat_p(plane_p * line_p) [n_rows, sum c] of the samples idx[:n_rows]
    of pts [..., 3] (all samples when idx is None): the gather of VmDecomposedTensor.get_color
    before basis_matrix_color (SimpleTensoRF09.py:1241-1263).  With ``count`` (device int32 [1], e.g.
    threshold_compact's) only the rows below it are gathered, n_rows being the capacity; the rows
    past it are left unwritten, and no host synchronisation is needed."""
    p = dev_tensor(pts.detach(), 'pts').reshape(-1, 3)
    idx, n_rows, count = _rows_and_count(p, idx, n_rows, count)
    return _VmAppearance.apply(p, idx, n_rows, count, tuple(float(x) for x in bbox), *planes, *lines)


def _cp_struct(lines):
    """struct srf_cp_factors of the three CP lines [1, R, L_i, 1] (plus the views it points into)."""
    f = _lib.CpFactors()
    keep = []
    rank = None
    for i in range(3):
        L = _texel_view(lines[i], f'line {i}')
        if L.shape[1] != 1 or (rank is not None and L.shape[2] != rank):
            raise _lib.SrfError(f'line {i}: expected [1, {rank or "R"}, L, 1], got {tuple(lines[i].shape)}')
        rank = L.shape[2]
        f.line_len[i] = L.shape[0]
        f.lines[i] = L.data_ptr()
        keep.append(L)
    f.n_comp = rank
    return f, keep


def _cp_grad_buffers(lines, needs):
    """Zeroed texel-major gradient buffers of the three lines."""
    bufs = []
    for i in range(3):
        _, c, h, w = lines[i].shape
        bufs.append(torch.zeros((1, h, w, c), dtype=torch.float32, device=lines[i].device))
    arr = (C.c_void_p * 3)(*[g.data_ptr() for g in bufs])
    return arr, [g.permute(0, 3, 1, 2) if needs[i] else None for i, g in enumerate(bufs)]


class _CpDensity(torch.autograd.Function):
    @staticmethod
    def forward(ctx, pts, bbox, alpha, predictor, offset, *lines):
        f, keep = _cp_struct(lines)
        n = pts.shape[0]
        dev = pts.device
        sigma = torch.empty((n,), dtype=torch.float32, device=dev)
        feat = torch.empty((n,), dtype=torch.float32, device=dev)
        valid = torch.empty((n,), dtype=torch.uint8, device=dev)
        g = alpha.struct() if alpha is not None else None
        _lib.call('srf_cp_density_fwd', C.byref(f), _bbox_arg(bbox), C.byref(g) if g is not None else None, n,
                  _lib.ptr(pts), predictor, float(offset), _lib.ptr(sigma), _lib.ptr(feat), _lib.ptr(valid),
                  _lib.stream_ptr(dev))
        ctx.save_for_backward(pts, feat, valid, *lines)
        ctx.cfg = (bbox, predictor, float(offset))
        ctx.mark_non_differentiable(valid)
        return sigma, valid

    @staticmethod
    def backward(ctx, g_sigma, _g_valid):
        pts, feat, valid, *lines = ctx.saved_tensors
        bbox, predictor, offset = ctx.cfg
        needs = ctx.needs_input_grad[5:]
        if not any(needs) or g_sigma is None:
            return (None,) * 8
        f, keep = _cp_struct(lines)
        dl, grads = _cp_grad_buffers(lines, needs)
        _lib.call('srf_cp_density_bwd', C.byref(f), dl, _bbox_arg(bbox), pts.shape[0], _lib.ptr(pts),
                  _lib.ptr(valid), _lib.ptr(feat), predictor, offset, _lib.ptr(g_sigma.contiguous()),
                  _lib.stream_ptr(pts.device))
        return (None, None, None, None, None, *grads)


def cp_density(pts: torch.Tensor, lines, bbox, *, alpha: AlphaMask | None = None, predictor: int = RELU,
               offset: float = 0.0):
    """vm_density for the CP decomposition: sigma and validity of pts [..., 3] with the density
    feature sum_r line_0[r] line_1[r] line_2[r] (CpDecomposedTensor.get_volume_density,
    SimpleTensoRF09.py:1043-1062).  Gradients flow to the three lines."""
    shape = pts.shape[:-1]
    p = dev_tensor(pts.detach(), 'pts').reshape(-1, 3)
    sigma, valid = _CpDensity.apply(p, tuple(float(x) for x in bbox), alpha, int(predictor), float(offset), *lines)
    return sigma.view(shape), valid.view(shape)


class _CpAppearance(torch.autograd.Function):
    @staticmethod
    def forward(ctx, pts, idx, n_rows, count, bbox, *lines):
        f, keep = _cp_struct(lines)
        out = torch.empty((n_rows, f.n_comp), dtype=torch.float32, device=pts.device)
        _lib.call('srf_cp_appearance_fwd', C.byref(f), _bbox_arg(bbox), n_rows, _lib.ptr(count), _lib.ptr(idx),
                  _lib.ptr(pts), _lib.ptr(out), _lib.stream_ptr(pts.device))
        ctx.save_for_backward(pts, idx, count, *lines)
        ctx.cfg = (n_rows, bbox)
        return out

    @staticmethod
    def backward(ctx, g_out):
        pts, idx, count, *lines = ctx.saved_tensors
        n_rows, bbox = ctx.cfg
        needs = ctx.needs_input_grad[5:]
        if not any(needs):
            return (None,) * 8
        f, keep = _cp_struct(lines)
        dl, grads = _cp_grad_buffers(lines, needs)
        g = g_out.contiguous()
        if g.data_ptr() % 16:
            g = g.clone()
        _lib.call('srf_cp_appearance_bwd', C.byref(f), dl, _bbox_arg(bbox), n_rows, _lib.ptr(count), _lib.ptr(idx),
                  _lib.ptr(pts), _lib.ptr(g), _lib.stream_ptr(pts.device))
        return (None, None, None, None, None, *grads)


def cp_appearance(pts: torch.Tensor, lines, bbox, idx: torch.Tensor | None = None, n_rows: int | None = None,
                  count: torch.Tensor | None = None):
    """CP appearance features [n_rows, R] (per-component products of the three lines) of the
    samples idx[:n_rows] of pts [..., 3] (all samples when idx is None): the gather of
    CpDecomposedTensor.get_color before basis_matrix_color (SimpleTensoRF09.py:1064-1078).  ``count``
    as for vm_appearance."""
    p = dev_tensor(pts.detach(), 'pts').reshape(-1, 3)
    idx, n_rows, count = _rows_and_count(p, idx, n_rows, count)
    return _CpAppearance.apply(p, idx, n_rows, count, tuple(float(x) for x in bbox), *lines)


def threshold_compact(values: torch.Tensor, threshold: float):
    """Ascending int32 indices of values.flatten() > threshold and their device count.

    Returns (idx [n] int32, count [1] int32); the first count entries of idx are valid.  The
    surface test of SimpleTensoRF09.py:726 with the boolean indexing of :1249.
    """
    v = dev_tensor(values.detach(), 'values').reshape(-1)
    n = v.shape[0]
    idx = torch.empty((n,), dtype=torch.int32, device=v.device)
    count = torch.empty((1,), dtype=torch.int32, device=v.device)
    ws = torch.empty((_lib.lib().srf_compact_workspace_bytes(n),), dtype=torch.uint8, device=v.device)
    _lib.call('srf_threshold_compact', n, _lib.ptr(v), float(threshold), _lib.ptr(idx), _lib.ptr(count),
              _lib.ptr(ws), _lib.stream_ptr(v.device))
    return idx, count


def aabb_samples(rays_o: torch.Tensor, rays_d: torch.Tensor, bbox, near: float, far: float, step: float,
                 n_samples: int, u: torch.Tensor | None = None) -> torch.Tensor:
    """Non-NDC coarse depths of SimpleTensoRF.get_z_vals_coarse (SimpleTensoRF09.py:382-400):
    z [N, n_samples] = t_min + step * (arange(n_samples) + u) with u [N] (or 0)."""
    n = rays_o.shape[0]
    rays_o = dev_tensor(rays_o, 'rays_o', (n, 3))
    rays_d = dev_tensor(rays_d, 'rays_d', (n, 3))
    if u is not None:
        u = dev_tensor(u, 'u', (n,))
    z = torch.empty((n, n_samples), dtype=torch.float32, device=rays_o.device)
    _lib.call('srf_aabb_samples', n, n_samples, _lib.ptr(rays_o), _lib.ptr(rays_d), _bbox_arg(bbox), float(near),
              float(far), float(step), _lib.ptr(u), _lib.ptr(z), _lib.stream_ptr(rays_o.device))
    return z
