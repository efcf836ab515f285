"""Oracle: Simple-TensoRF hot path restated in plain PyTorch (TEST INFRASTRUCTURE ONLY).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may use this
module.  Functions follow src/models/SimpleTensoRF09.py (cited per function).
"""
from __future__ import annotations

import math

import torch
import torch.nn.functional as F

from .nerf import positional_encoding

# Axis assignment of the vector-matrix decomposition (SimpleTensoRF09.py:1131-1132).
MATRIX_AXES = [[0, 1], [0, 2], [1, 2]]
VECTOR_AXES = [2, 1, 0]


def compute_resolution_in_voxels(num_voxels: int, bounding_box: torch.Tensor) -> torch.Tensor:
    """SimpleTensoRF09.py:626-638."""
    xyz_min, xyz_max = bounding_box
    voxel_size = ((xyz_max - xyz_min).prod() / num_voxels).pow(1 / 3)
    return ((xyz_max - xyz_min) / voxel_size).long()


def compute_num_samples(resolution: torch.Tensor, num_voxels_per_sample: float, num_samples_max) -> int:
    """SimpleTensoRF09.py:640-650."""
    n = int((torch.linalg.norm(resolution.float()) / num_voxels_per_sample).round().long())
    return int(min(num_samples_max, n))


def step_size(bounding_box: torch.Tensor, resolution: torch.Tensor, num_voxels_per_sample: float) -> float:
    """SimpleTensoRF09.py:658-663."""
    voxel_length = (bounding_box[1] - bounding_box[0]) / (resolution - 1)
    return float(torch.mean(voxel_length) * num_voxels_per_sample)


def normalize_points(pts: torch.Tensor, bounding_box: torch.Tensor) -> torch.Tensor:
    """SimpleTensoRF09.py:763-765."""
    return ((pts - bounding_box[0]) / (bounding_box[1] - bounding_box[0])) * 2 - 1


def bbox_validity(pts: torch.Tensor, bounding_box: torch.Tensor) -> torch.Tensor:
    """SimpleTensoRF09.py:705."""
    return ((bounding_box[0] <= pts) & (pts <= bounding_box[1])).all(dim=-1)


def alpha_mask_sample(alpha_volume: torch.Tensor, bounding_box: torch.Tensor, pts: torch.Tensor) -> torch.Tensor:
    """Trilinear occupancy lookup. AlphaGridMask.sample_alpha, SimpleTensoRF09.py:1342-1349.
    alpha_volume: [D,H,W] (z,y,x order, as built at :860-868)."""
    vol = alpha_volume.view(1, 1, *alpha_volume.shape[-3:]).to(pts.dtype)
    pn = normalize_points(pts, bounding_box)
    return F.grid_sample(vol, pn.view(1, -1, 1, 1, 3), align_corners=True).view(-1)


def _plane_line_coords(pts_n: torch.Tensor):
    """Coordinates as built at SimpleTensoRF09.py:1224-1228."""
    cp = torch.stack((pts_n[..., MATRIX_AXES[0]], pts_n[..., MATRIX_AXES[1]],
                      pts_n[..., MATRIX_AXES[2]])).detach().view(3, -1, 1, 2)
    cl = torch.stack((pts_n[..., VECTOR_AXES[0]], pts_n[..., VECTOR_AXES[1]], pts_n[..., VECTOR_AXES[2]]))
    cl = torch.stack((torch.zeros_like(cl), cl), dim=-1).detach().view(3, -1, 1, 2)
    return cp, cl


def vm_density_features(planes, lines, pts_n: torch.Tensor) -> torch.Tensor:
    """sigma features = sum_p sum_c plane_p(x,y) * line_p(z).
    VmDecomposedTensor.get_volume_density, SimpleTensoRF09.py:1214-1239 (before the
    density predictor).  planes[p]: [1,c_p,H,W]; lines[p]: [1,c_p,L,1]; pts_n [M,3]."""
    cp, cl = _plane_line_coords(pts_n)
    m = pts_n.shape[0]
    feat = torch.zeros((m,), dtype=pts_n.dtype, device=pts_n.device)
    for p in range(len(planes)):
        pc = F.grid_sample(planes[p], cp[[p]], align_corners=True).view(-1, m)
        lc = F.grid_sample(lines[p], cl[[p]], align_corners=True).view(-1, m)
        feat = feat + torch.sum(pc * lc, dim=0)
    return feat


def vm_appearance_features(planes, lines, pts_n: torch.Tensor) -> torch.Tensor:
    """[M, sum_p c_p] = cat_p(plane_p * line_p).  VmDecomposedTensor.get_color,
    SimpleTensoRF09.py:1241-1263 (before basis_matrix_color)."""
    cp, cl = _plane_line_coords(pts_n)
    m = pts_n.shape[0]
    pcs, lcs = [], []
    for p in range(len(planes)):
        pcs.append(F.grid_sample(planes[p], cp[[p]], align_corners=True).view(-1, m))
        lcs.append(F.grid_sample(lines[p], cl[[p]], align_corners=True).view(-1, m))
    return (torch.cat(pcs) * torch.cat(lcs)).T


def _line_coords(pts_n: torch.Tensor):
    """Line coordinates of the CP decomposition, SimpleTensoRF09.py:1051-1052 (same as :1226-1227)."""
    cl = torch.stack((pts_n[..., VECTOR_AXES[0]], pts_n[..., VECTOR_AXES[1]], pts_n[..., VECTOR_AXES[2]]))
    return torch.stack((torch.zeros_like(cl), cl), dim=-1).detach().view(3, -1, 1, 2)


def cp_component_products(lines, pts_n: torch.Tensor) -> torch.Tensor:
    """[R, M] = (line_0 * line_1) * line_2 per component, the three lines of rank R sampled at
    pts[2], pts[1], pts[0].  CpDecomposedTensor.get_volume_density / get_color,
    SimpleTensoRF09.py:1051-1056 and :1072-1077; lines[i]: [1,R,L_i,1] (built at :989-996)."""
    cl = _line_coords(pts_n)
    m = pts_n.shape[0]
    prod = F.grid_sample(lines[0], cl[[0]], align_corners=True).view(-1, m)
    prod = prod * F.grid_sample(lines[1], cl[[1]], align_corners=True).view(-1, m)
    prod = prod * F.grid_sample(lines[2], cl[[2]], align_corners=True).view(-1, m)
    return prod


def cp_density_features(lines, pts_n: torch.Tensor) -> torch.Tensor:
    """sigma features = sum_r of the component products, SimpleTensoRF09.py:1057 (before the
    density predictor)."""
    return torch.sum(cp_component_products(lines, pts_n), dim=0)


def cp_appearance_features(lines, pts_n: torch.Tensor) -> torch.Tensor:
    """[M, R] component products, the input of basis_matrix_color at SimpleTensoRF09.py:1078."""
    return cp_component_products(lines, pts_n).T


def bilinear_align_corners(img: torch.Tensor, x: torch.Tensor, y: torch.Tensor) -> torch.Tensor:
    """Closed-form bilinear sampling with align_corners=True and zero padding (the
    textbook definition grid_sample implements); img [C,H,W], x,y [M] in [-1,1].
    Used to cross-check grid_sample in the tests."""
    C, H, W = img.shape
    ix = (x + 1) / 2 * (W - 1)
    iy = (y + 1) / 2 * (H - 1)
    x0 = torch.floor(ix)
    y0 = torch.floor(iy)
    out = torch.zeros((C, x.shape[0]), dtype=img.dtype, device=img.device)
    for dx in (0, 1):
        for dy in (0, 1):
            xi = x0 + dx
            yi = y0 + dy
            wx = (ix - x0) if dx else (1 - (ix - x0))
            wy = (iy - y0) if dy else (1 - (iy - y0))
            ok = (xi >= 0) & (xi <= W - 1) & (yi >= 0) & (yi <= H - 1)
            xc = xi.clamp(0, W - 1).long()
            yc = yi.clamp(0, H - 1).long()
            out += img[:, yc, xc] * (wx * wy * ok.to(img.dtype))
    return out


def density_predictor(feat: torch.Tensor, kind: str = 'ReLU', offset: float = -10.0) -> torch.Tensor:
    """SimpleTensoRF09.py:671-679."""
    if kind == 'ReLU':
        return F.relu(feat)
    if kind == 'SoftPlus':
        return F.softplus(feat + offset)
    raise NotImplementedError(kind)


def color_mlp_forward(params: dict, features: torch.Tensor, view_dirs: torch.Tensor | None,
                      features_pe_degree: int = 0, views_pe_degree: int = 0) -> torch.Tensor:
    """MlpFeaturesColorPredictor.forward, SimpleTensoRF09.py:1411-1421 (layers :1389-1393).
    params: mlp.0.{weight,bias}, mlp.2.*, mlp.4.*."""
    x = positional_encoding(features, features_pe_degree)
    if view_dirs is not None:
        x = torch.cat([x, positional_encoding(view_dirs, views_pe_degree)], dim=-1)
    h = F.relu(F.linear(x, params['mlp.0.weight'], params['mlp.0.bias']))
    h = F.relu(F.linear(h, params['mlp.2.weight'], params['mlp.2.bias']))
    return torch.sigmoid(F.linear(h, params['mlp.4.weight'], params['mlp.4.bias']))


def aabb_z_vals(rays_o: torch.Tensor, rays_d: torch.Tensor, bounding_box: torch.Tensor, near: float, far: float,
                step: float, n_samples: int, u: torch.Tensor | None = None) -> torch.Tensor:
    """Non-NDC coarse depths, SimpleTensoRF09.py:382-400: ray entry into the box, clamped to
    [near, far], plus step * (arange(S) + u) with one u per ray."""
    vec = torch.where(rays_d == 0, torch.full_like(rays_d, 1e-6), rays_d)
    rate_a = (bounding_box[1] - rays_o) / vec
    rate_b = (bounding_box[0] - rays_o) / vec
    t_min = torch.minimum(rate_a, rate_b).amax(-1).clamp(min=near, max=far)
    rng = torch.arange(n_samples, dtype=rays_o.dtype, device=rays_o.device)[None]
    if u is not None:
        rng = rng.repeat(rays_d.shape[0], 1)
        rng = rng + u.to(rays_o.dtype)[:, None]
    st = torch.tensor(step, dtype=rays_o.dtype, device=rays_o.device) * rng
    return t_min[..., None] + st


def tensorf_z_vals_ndc(n_rays: int, n_samples: int, near: float, far: float, u=None, dtype=torch.float32):
    """NDC coarse depths of SimpleTensoRF09.py:364-381 (same construction as NeRF)."""
    from .nerf import stratified_z_vals
    return stratified_z_vals(n_rays, n_samples, near, far, False, u=u, dtype=dtype)


def tensorf_render(model: dict, pts: torch.Tensor, z_vals_ndc: torch.Tensor, rays_o, rays_d, rays_d_ndc,
                   view_dirs: torch.Tensor, *, alpha_volume=None, distance_scale=25.0, weight_threshold=1e-4,
                   density_kind='ReLU', density_offset=-10.0, white_bkgd=False, features_pe_degree=0,
                   views_pe_degree=0, ndc=True):
    """LowRankTensor.forward in NDC, SimpleTensoRF09.py:701-761, for the VM decomposition
    (:1214-1263) or, when model['decomposition'] == 'CandecompParafac', the CP one (:1043-1089).

    model: dict(bbox [2,3], planes_density, lines_density, planes_color, lines_color (no planes
    for CP), basis [27, sum c], color_mlp params).  view_dirs [N,3] (expanded per sample as at
    :735).  With ndc=False, z_vals_ndc holds the metric depths and the non-NDC branch of :775-778
    applies (delta from the metric depths, infinite last interval).  Returns the compositing dict
    plus raw sigma/rgb.
    """
    cp = model.get('decomposition', 'VectorMatrix') == 'CandecompParafac'
    if cp:
        density_features = lambda q: cp_density_features(model['lines_density'], q)
        appearance_features = lambda q: cp_appearance_features(model['lines_color'], q)
    else:
        density_features = lambda q: vm_density_features(model['planes_density'], model['lines_density'], q)
        appearance_features = lambda q: vm_appearance_features(model['planes_color'], model['lines_color'], q)
    from .nerf import volume_rendering
    bbox = model['bbox']
    valid = bbox_validity(pts, bbox)
    if alpha_volume is not None:
        alphas = alpha_mask_sample(alpha_volume, bbox, pts[valid])
        valid[valid.clone()] &= alphas > 0
    pts_n = normalize_points(pts, bbox)
    sigma = torch.zeros(pts.shape[:-1], dtype=pts.dtype, device=pts.device)
    if valid.any():
        feat = density_features(pts_n[valid])
        sigma[valid] = density_predictor(feat, density_kind, density_offset)
    zero_rgb = torch.zeros(*pts.shape[:-1], 3, dtype=pts.dtype, device=pts.device)
    if ndc:
        out = volume_rendering(sigma, zero_rgb, z_vals_ndc=z_vals_ndc, rays_o=rays_o, rays_d=rays_d,
                               rays_d_ndc=rays_d_ndc, ndc=True, distance_scale=distance_scale)
    else:
        out = volume_rendering(sigma, zero_rgb, z_vals=z_vals_ndc, rays_o=rays_o, rays_d=rays_d, ndc=False,
                               distance_scale=distance_scale)
    surface = out['weights'] > weight_threshold
    rgb = torch.zeros(*pts.shape[:-1], 3, dtype=pts.dtype, device=pts.device)
    if surface.any():
        app = appearance_features(pts_n[surface])
        feats = F.linear(app, model['basis'])
        vd = view_dirs[:, None].expand(pts.shape)[surface]
        rgb[surface] = color_mlp_forward(model['color_mlp'], feats, vd, features_pe_degree, views_pe_degree)
    rgb_map = torch.sum(out['weights'][..., None] * rgb, dim=-2)
    if white_bkgd:
        rgb_map = rgb_map + (1 - out['acc'][..., None])
    out['rgb'] = rgb_map
    out['raw_sigma'] = sigma[..., None]
    out['raw_rgb'] = rgb
    return out


def tensorf_render_ragged(model: dict, pts: torch.Tensor, z_vals: torch.Tensor, ray_offsets, rays_o, rays_d,
                          view_dirs, *, alpha_volume=None, distance_scale=25.0, weight_threshold=1e-4,
                          density_kind='ReLU', density_offset=-10.0, white_bkgd=False, features_pe_degree=0,
                          views_pe_degree=0):
    """LowRankTensor.forward (non-NDC, SimpleTensoRF09.py:701-761, :767-819) over a ragged batch:
    ray r owns the flat samples [ray_offsets[r], ray_offsets[r+1]) of pts [P,3] / z_vals [P]
    (metric depths).  Per ray this is the reference's forward on a batch of that one ray: the
    validity, density, surface test, appearance gather, basis and colour MLP are per sample (so
    they run over all samples at once), and the compositing (:775-803, infinite last interval) runs
    ray by ray.  Returns per-ray rgb/acc/depth/depth_var and per-sample weights/raw_sigma/raw_rgb."""
    from .nerf import volume_rendering
    cp = model.get('decomposition', 'VectorMatrix') == 'CandecompParafac'
    bbox = model['bbox']
    valid = bbox_validity(pts, bbox)
    if alpha_volume is not None:
        alphas = alpha_mask_sample(alpha_volume, bbox, pts[valid])
        valid[valid.clone()] &= alphas > 0
    pts_n = normalize_points(pts, bbox)
    sigma = torch.zeros(pts.shape[:-1], dtype=pts.dtype, device=pts.device)
    if valid.any():
        if cp:
            feat = cp_density_features(model['lines_density'], pts_n[valid])
        else:
            feat = vm_density_features(model['planes_density'], model['lines_density'], pts_n[valid])
        sigma[valid] = density_predictor(feat, density_kind, density_offset)
    off = [int(x) for x in ray_offsets]
    n = len(off) - 1
    counts = torch.tensor([off[r + 1] - off[r] for r in range(n)], device=pts.device)
    ray_of = torch.repeat_interleave(torch.arange(n, device=pts.device), counts)
    per = {'acc': [], 'depth': [], 'depth_var': [], 'weights': []}
    for r in range(n):
        a, b = off[r], off[r + 1]
        o = volume_rendering(sigma[a:b][None], torch.zeros(1, b - a, 3, dtype=pts.dtype, device=pts.device),
                             z_vals=z_vals[a:b][None], rays_o=rays_o[r:r + 1], rays_d=rays_d[r:r + 1], ndc=False,
                             distance_scale=distance_scale)
        for k in per:
            per[k].append(o[k][0] if k == 'weights' else o[k])
    res = {k: torch.cat(v) for k, v in per.items()}
    weights = res['weights']
    surface = weights > weight_threshold
    rgb = torch.zeros(*pts.shape[:-1], 3, dtype=pts.dtype, device=pts.device)
    if surface.any():
        if cp:
            app = cp_appearance_features(model['lines_color'], pts_n[surface])
        else:
            app = vm_appearance_features(model['planes_color'], model['lines_color'], pts_n[surface])
        feats = F.linear(app, model['basis'])
        rgb[surface] = color_mlp_forward(model['color_mlp'], feats, view_dirs[ray_of][surface], features_pe_degree,
                                         views_pe_degree)
    rgb_map = torch.zeros(n, 3, dtype=pts.dtype, device=pts.device).index_add(0, ray_of, weights[:, None] * rgb)
    if white_bkgd:
        rgb_map = rgb_map + (1 - res['acc'][..., None])
    res['rgb'] = rgb_map
    res['raw_sigma'] = sigma[..., None]
    res['raw_rgb'] = rgb
    return res
