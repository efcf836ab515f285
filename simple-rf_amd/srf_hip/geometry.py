"""Ray generation and NDC warping (src/utils/CommonUtils04.py:73-149).

rays_fused runs get_rays_tr, get_view_dirs_tr and get_ndc_rays_tr as one HIP kernel (srf_rays) and
is what the models use when the camera carries no gradient.  The PyTorch functions restate the
reference op for op; they are the differentiable path for a model that learns its camera.
"""
from __future__ import annotations

import torch


def get_rays_tr(pixel_id: torch.Tensor, intrinsic: torch.Tensor, extrinsics: torch.Tensor, mip_nerf_used: bool,
                nerf_synthetic: bool = False, intrinsic_inv: torch.Tensor | None = None):
    """src/utils/CommonUtils04.py:73-95.  intrinsic_inv: the inverse intrinsics per ray when the caller
    has them (inverted once per frame); otherwise each ray's matrix is inverted here."""
    if nerf_synthetic:
        p = torch.eye(4).to(extrinsics)
        p[1, 1] = -1
        p[2, 2] = -1
        extrinsics = torch.matmul(torch.matmul(p, extrinsics), p)
    pixel_id = pixel_id.float()
    x, y = pixel_id[:, 1], pixel_id[:, 2]
    if mip_nerf_used:
        x = x + 0.5
        y = y + 0.5
    points_homo = torch.stack([x, y, torch.ones_like(x)], dim=1)
    if intrinsic_inv is None:
        # the same LU inverse as torch.linalg.inv, without its host-synchronising singularity check
        # (keeps the step capturable in a HIP graph)
        intrinsic_inv = torch.linalg.inv_ex(intrinsic)[0]
    # K^-1 [x, y, 1] per ray as one elementwise product and sum (no batched 3 x 3 GEMM)
    dirs = torch.sum(intrinsic_inv * points_homo[:, None, :], dim=-1)
    dirs[:, 1:] *= -1
    if nerf_synthetic:
        dirs[:, 1:] *= -1
    rays_d = torch.sum(dirs[:, None, :] * extrinsics[:, :3, :3], dim=-1)
    rays_o = extrinsics[:, :3, -1]
    return rays_o, rays_d


def get_ndc_rays_tr(rays_o, rays_d, resolution, intrinsic, near):
    """src/utils/CommonUtils04.py:120-138."""
    h, w = resolution
    fx, fy = intrinsic[:, 0, 0], intrinsic[:, 1, 1]
    t = -(near + rays_o[:, 2]) / rays_d[:, 2]
    rays_o = rays_o + t[:, None] * rays_d
    o0 = -1. / (w / (2. * fx)) * rays_o[:, 0] / rays_o[:, 2]
    o1 = -1. / (h / (2. * fy)) * rays_o[:, 1] / rays_o[:, 2]
    o2 = 1. + 2. * near / rays_o[:, 2]
    d0 = -1. / (w / (2. * fx)) * (rays_d[:, 0] / rays_d[:, 2] - rays_o[:, 0] / rays_o[:, 2])
    d1 = -1. / (h / (2. * fy)) * (rays_d[:, 1] / rays_d[:, 2] - rays_o[:, 1] / rays_o[:, 2])
    d2 = -2. * near / rays_o[:, 2]
    return torch.stack([o0, o1, o2], -1), torch.stack([d0, d1, d2], -1)


def get_view_dirs_tr(rays_d):
    """src/utils/CommonUtils04.py:147-149."""
    return rays_d / torch.linalg.norm(rays_d, ord=2, dim=-1, keepdim=True)


def rays_fused(pixel_id: torch.Tensor, intrinsics: torch.Tensor, extrinsics: torch.Tensor, *, mip_nerf_used: bool,
               ndc: bool, resolution=None, near: float = 0.0):
    """get_rays_tr, get_view_dirs_tr and (ndc) get_ndc_rays_tr in one HIP kernel (srf_rays).

    pixel_id [N,3] (frame, x, y); intrinsics [F,3,3] and extrinsics (camera-to-world) [F,4,4] per
    frame, indexed by pixel_id[:, 0].  Returns rays_o, rays_d, view_dirs and rays_o_ndc,
    rays_d_ndc (None without ndc).  Not differentiable: callers that learn the camera use the
    PyTorch functions above."""
    from . import _lib
    from ._checks import dev_tensor
    n = pixel_id.shape[0]
    dev = pixel_id.device
    pid = dev_tensor(pixel_id.to(torch.int32), 'pixel_id', (n, 3), torch.int32)
    n_frames = intrinsics.shape[0]
    K = dev_tensor(intrinsics.detach().float(), 'intrinsics', (n_frames, 3, 3))
    E = dev_tensor(extrinsics.detach().float(), 'extrinsics', (n_frames, 4, 4))
    f = dict(dtype=torch.float32, device=dev)
    rays_o, rays_d, view_dirs = (torch.empty((n, 3), **f) for _ in range(3))
    o_ndc = torch.empty((n, 3), **f) if ndc else None
    d_ndc = torch.empty((n, 3), **f) if ndc else None
    h, w = resolution if ndc else (0, 0)
    _lib.call('srf_rays', n, _lib.ptr(pid), n_frames, _lib.ptr(K), _lib.ptr(E), int(mip_nerf_used), int(ndc),
              int(h), int(w), float(near), _lib.ptr(rays_o), _lib.ptr(rays_d), _lib.ptr(view_dirs),
              _lib.ptr(o_ndc), _lib.ptr(d_ndc), _lib.stream_ptr(dev))
    return rays_o, rays_d, view_dirs, o_ndc, d_ndc
