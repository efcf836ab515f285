"""Ray sampling ops on the HIP kernels (stratified, inverse CDF, hierarchical merge).

Mirrors SimpleNeRF.get_z_vals_coarse / sample_pdf / get_z_vals_fine
(src/models/SimpleNeRF17.py:330-417) with the reference's CPU random draws replaced by
explicit ``u`` tensors (None selects the deterministic, perturb=False path).
No gradient flows through sampling (the reference detaches z_samples at :368).
"""
from __future__ import annotations

import torch

from . import _lib
from ._checks import dev_tensor, opt_tensor


def stratified_z_vals(n_rays: int, n_samples: int, near: float, far: float, *, lindisp: bool = False,
                      u: torch.Tensor | None = None, device=None) -> torch.Tensor:
    """Coarse depths [n_rays, n_samples] (SimpleNeRF17.py:330-358)."""
    if u is not None:
        u = dev_tensor(u, 'u', (n_rays, n_samples))
        device = u.device
    device = torch.device(device if device is not None else 'cuda')
    z = torch.empty((n_rays, n_samples), dtype=torch.float32, device=device)
    _lib.call('srf_stratified_samples', n_rays, n_samples, float(near), float(far), int(bool(lindisp)),
              _lib.ptr(u), _lib.ptr(z), _lib.stream_ptr(device))
    return z


def sample_pdf(bins: torch.Tensor, weights: torch.Tensor | None, n_samples: int, *, u=None, cdf=None,
               return_inds: bool = False, return_cdf: bool = False):
    """Inverse-CDF sampling (SimpleNeRF17.py:384-417).  ``u`` None means det=True."""
    bins = dev_tensor(bins, 'bins')
    n_rays, n_bins = bins.shape
    weights = opt_tensor(weights, 'weights', (n_rays, n_bins - 1))
    cdf = opt_tensor(cdf, 'cdf', (n_rays, n_bins))
    u = opt_tensor(u, 'u', (n_rays, n_samples))
    samples = torch.empty((n_rays, n_samples), dtype=torch.float32, device=bins.device)
    inds = torch.empty((n_rays, n_samples), dtype=torch.int64, device=bins.device) if return_inds else None
    cdf_out = torch.empty((n_rays, n_bins), dtype=torch.float32, device=bins.device) if return_cdf else None
    _lib.call('srf_sample_pdf', n_rays, n_bins, n_samples, _lib.ptr(bins), _lib.ptr(weights), _lib.ptr(cdf),
              _lib.ptr(u), _lib.ptr(samples), _lib.ptr(inds), _lib.ptr(cdf_out), _lib.stream_ptr(bins.device))
    out = [samples]
    if return_inds:
        out.append(inds)
    if return_cdf:
        out.append(cdf_out)
    return out[0] if len(out) == 1 else tuple(out)


def fine_z_vals(z_coarse: torch.Tensor, weights_coarse: torch.Tensor, n_fine: int, *, u=None,
                return_fine: bool = False):
    """sort(cat(z_coarse, sample_pdf(mid(z_coarse), w[...,1:-1]))) (SimpleNeRF17.py:360-371)."""
    z_coarse = dev_tensor(z_coarse.detach(), 'z_coarse')
    n_rays, n_coarse = z_coarse.shape
    weights_coarse = dev_tensor(weights_coarse.detach(), 'weights_coarse', (n_rays, n_coarse))
    u = opt_tensor(u, 'u', (n_rays, n_fine))
    z_merged = torch.empty((n_rays, n_coarse + n_fine), dtype=torch.float32, device=z_coarse.device)
    z_fine = torch.empty((n_rays, n_fine), dtype=torch.float32, device=z_coarse.device) if return_fine else None
    _lib.call('srf_fine_z_vals', n_rays, n_coarse, n_fine, _lib.ptr(z_coarse), _lib.ptr(weights_coarse),
              _lib.ptr(u), _lib.ptr(z_merged), _lib.ptr(z_fine), _lib.stream_ptr(z_coarse.device))
    return (z_merged, z_fine) if return_fine else z_merged
