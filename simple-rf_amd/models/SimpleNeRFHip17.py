"""Drop-in replacement for the Simple-NeRF model on the MI355X HIP kernels.

Selected by the reference's plugin loader (src/models/ModelFactory02.py:10-22) with
``configs['model']['name'] = 'SimpleNeRFHip17'`` (class ``SimpleNeRFHip``).  The class surface,
config keys, output-dict keys and state_dict names/shapes are those of SimpleNeRF
(src/models/SimpleNeRF17.py:14-578, MLP :616-785, camera learners :788-912), so checkpoints
interchange and Trainer10/Tester07 drive it unchanged.  The hot path runs on the HIP kernels:

  get_z_vals_coarse  -> srf_stratified_samples        (SimpleNeRF17.py:330-358)
  get_z_vals_fine    -> srf_fine_z_vals               (:360-371, sample_pdf :384-417)
  run_network / MLP  -> srf_mlp_forward / _backward   (:419-484, :696-785)
  volume_rendering   -> srf_composite_fwd / _bwd      (:486-558)

Random draws the reference makes on the CPU generator inside the path (t_rand :355, u :397,
sigma noise :740) are drawn on the GPU (torch.rand / torch.randn on the ray device, from
``self.generator`` when set) and passed to the kernels as explicit inputs.  Extra config key:
``configs['model']['mlp_precision']`` in {'bf16', 'bf16_split', 'fp32'} (default 'bf16'; 'bf16_split'
rounds the forward like FP32 and the backward with two BF16 planes, on BF16 MFMAs).
"""
from __future__ import annotations

import numpy
import torch
from torch.nn import ModuleDict, ModuleList

from srf_hip import compositing, geometry, mlp as hip_mlp, sampling
from srf_hip._lib import SrfError


class SimpleNeRFHip(torch.nn.Module):
    def __init__(self, configs: dict, model_configs: dict):
        super().__init__()
        self.configs = configs
        self.model_configs = model_configs
        self.ndc = self.configs['data_loader']['ndc']
        self.coarse_model_needed = 'coarse_model' in self.configs['model']
        self.fine_model_needed = 'fine_model' in self.configs['model']
        self.predict_visibility = (self.coarse_model_needed and self.configs['model']['coarse_model']['predict_visibility']) or \
                                  (self.fine_model_needed and self.configs['model']['fine_model']['predict_visibility'])
        if self.predict_visibility:
            raise NotImplementedError('predict_visibility is not supported by the HIP MLP kernels')
        self.generator = None
        self.coarse_model = None
        self.fine_model = None
        self.intrinsics_learner = None
        self.extrinsics_learner = None
        self.augmentations_needed = 'augmentations' in configs['model']
        if self.augmentations_needed:
            self.augmented_models = []
            self.augmented_models_nn = []
        self.build_nerf()

    def build_nerf(self):
        """src/models/SimpleNeRF17.py:36-75."""
        if self.coarse_model_needed:
            self.coarse_model = MLP('coarse_model', self.configs, self.configs['model']['coarse_model'], self.model_configs)
        if self.fine_model_needed:
            self.fine_model = MLP('fine_model', self.configs, self.configs['model']['fine_model'], self.model_configs)
        self.intrinsics_learner = IntrinsicsLearner(numpy.array(self.model_configs['intrinsics']),
                                                    learn_focal=self.configs['model']['learn_camera_focal_length'])
        self.extrinsics_learner = ExtrinsicsLearner(numpy.array(self.model_configs['extrinsics']),
                                                    learn_rotation=self.configs['model']['learn_camera_rotation'],
                                                    learn_translation=self.configs['model']['learn_camera_translation'])
        if self.augmentations_needed:
            for aug in self.configs['model']['augmentations']:
                nn = ModuleDict()
                coarse = fine = None
                if 'coarse_model' in aug:
                    coarse = MLP(f"{aug['name']}_coarse_model", self.configs, aug['coarse_model'], self.model_configs)
                    nn['coarse_model'] = coarse
                if 'fine_model' in aug:
                    fine = MLP(f"{aug['name']}_fine_model", self.configs, aug['fine_model'], self.model_configs)
                    nn['fine_model'] = fine
                self.augmented_models.append({'name': aug['name'], 'coarse_model': coarse, 'fine_model': fine})
                self.augmented_models_nn.append(nn)
            self.augmented_models_nn = ModuleList(self.augmented_models_nn)

    def get_trainable_parameters(self, optimizer_configs):
        """src/models/SimpleNeRF17.py:77-90."""
        params = []
        if self.coarse_model_needed:
            params.extend(self.coarse_model.get_trainable_parameters(optimizer_configs))
        if self.fine_model_needed:
            params.extend(self.fine_model.get_trainable_parameters(optimizer_configs))
        if self.augmentations_needed:
            for aug in self.augmented_models:
                for k in ('coarse_model', 'fine_model'):
                    if aug[k] is not None:
                        params.extend(aug[k].get_trainable_parameters(optimizer_configs))
        return params

    def rebuild_camera_params_learners(self, *, intrinsics: numpy.ndarray = None, extrinsics=None, device):
        """src/models/SimpleNeRF17.py:92-102."""
        if intrinsics is not None:
            self.intrinsics_learner = IntrinsicsLearner(intrinsics, self.configs['model']['learn_camera_focal_length']).to(device)
        if extrinsics is not None:
            self.extrinsics_learner = ExtrinsicsLearner(extrinsics, self.configs['model']['learn_camera_rotation'],
                                                        self.configs['model']['learn_camera_translation']).to(device)

    # ------------------------------------------------------------------ forward
    def forward(self, input_batch: dict, *, retraw: bool = False, sec_views_vis: bool = False, mode: str = None):
        """src/models/SimpleNeRF17.py:104-127."""
        input_batch = self.deep_dict_copy(input_batch)
        if 'common_data' in input_batch.keys():
            for key in input_batch['common_data'].keys():
                if isinstance(input_batch['common_data'][key], torch.Tensor):
                    input_batch['common_data'][key] = input_batch['common_data'][key][0]
        image_id = input_batch['pixel_id'][:, 0]
        intrinsics = self.intrinsics_learner(image_id)
        # the camera of a ray is a function of its frame: evaluate the learners and the intrinsics
        # inverse once per frame and gather them per ray (the reference evaluates them per ray)
        cams = torch.arange(self.extrinsics_learner.num_frames, device=image_id.device, dtype=image_id.dtype)
        cam_extrinsics = self.extrinsics_learner(cams)
        extrinsics = cam_extrinsics[image_id]
        n_frames = input_batch['num_frames']
        all_extrinsics = cam_extrinsics if n_frames == cams.shape[0] else \
            self.extrinsics_learner(torch.arange(n_frames, device=image_id.device, dtype=image_id.dtype))
        input_batch['intrinsics'] = intrinsics
        k_cams = torch.arange(self.intrinsics_learner.initial_intrinsics.shape[0], device=image_id.device,
                              dtype=image_id.dtype)
        # per-frame cameras (a tuple, so that batchify_rays does not slice them as per-ray data)
        input_batch['camera_frames'] = (self.intrinsics_learner(k_cams), cam_extrinsics)
        input_batch['extrinsics'] = extrinsics
        if mode == 'camera_params_only':
            out = {}
        else:
            out = self.render(input_batch, retraw=retraw or self.training, sec_views_vis=sec_views_vis or self.training,
                              mode=mode)
        out['intrinsics'] = intrinsics
        out['extrinsics'] = extrinsics
        out['extrinsics_all'] = all_extrinsics
        return out

    def render(self, input_dict: dict, *, retraw: bool, sec_views_vis: bool, mode: str):
        return self.batchify_rays(input_dict, retraw=retraw, sec_views_vis=sec_views_vis, mode=mode)

    def batchify_rays(self, input_dict: dict, *, retraw, sec_views_vis, mode):
        """src/models/SimpleNeRF17.py:133-157 (chunks of configs['model']['chunk'] rays)."""
        all_ret = {}
        num_rays = input_dict['pixel_id'].shape[0]
        chunk = self.configs['model']['chunk']
        for i in range(0, num_rays, chunk):
            sub = {}
            for key in input_dict:
                v = input_dict[key]
                if isinstance(v, (torch.Tensor, numpy.ndarray)) and v.shape[0] == num_rays:
                    sub[key] = v[i:i + chunk]
                else:
                    sub[key] = v
            ret = self.render_rays(sub, retraw=retraw, sec_views_vis=sec_views_vis, mode=mode)
            for k in ret:
                all_ret.setdefault(k, []).append(ret[k])
        return self.merge_mini_batch_data(all_ret)

    def render_rays(self, input_dict: dict, *, retraw, sec_views_vis, mode):
        """src/models/SimpleNeRF17.py:159-328."""
        ret = {}
        pixel_id = input_dict['pixel_id']
        intrinsics, extrinsics = input_dict['intrinsics'], input_dict['extrinsics']
        frames = input_dict.get('camera_frames')
        fused = frames is not None and mode != 'static_camera' and \
            not (torch.is_grad_enabled() and (frames[0].requires_grad or frames[1].requires_grad))
        rays_o_ndc = rays_d_ndc = fused_view_dirs = None
        if fused:
            # rays, view directions and the NDC warp in one kernel (the camera carries no gradient)
            rays_o, rays_d, fused_view_dirs, rays_o_ndc, rays_d_ndc = geometry.rays_fused(
                pixel_id, frames[0], frames[1], mip_nerf_used=False, ndc=self.ndc,
                resolution=self.model_configs['resolution'], near=self.model_configs['near'])
        else:
            intrinsics_inv = None if frames is None else \
                torch.linalg.inv_ex(frames[0])[0][pixel_id[:, 0].long()]
            rays_o, rays_d = geometry.get_rays_tr(pixel_id, intrinsics, extrinsics, mip_nerf_used=False,
                                                  intrinsic_inv=intrinsics_inv)
            if self.ndc:
                rays_o_ndc, rays_d_ndc = geometry.get_ndc_rays_tr(rays_o, rays_d, self.model_configs['resolution'],
                                                                  intrinsics, near=self.model_configs['near'])
        ret['rays_o'], ret['rays_d'] = rays_o, rays_d
        if self.ndc:
            ret['rays_o_ndc'], ret['rays_d_ndc'] = rays_o_ndc, rays_d_ndc
        if torch.is_grad_enabled() and (rays_o.requires_grad or rays_d.requires_grad):
            # a learned camera (learn_camera_* or camera optimisation at test time) differentiates the
            # loss through the positional encoding into the rays (src/models/SimpleNeRF17.py:831-842);
            # the fused kernels give no gradient to points or rays
            raise NotImplementedError('camera gradients through the rays are not supported by the HIP kernels')
        view_dirs = None
        if (self.coarse_model_needed and self.configs['model']['coarse_model']['use_view_dirs']) or \
                (self.fine_model_needed and self.configs['model']['fine_model']['use_view_dirs']):
            if fused:
                view_dirs = fused_view_dirs
            elif mode == 'static_camera':
                cd = input_dict['common_data']
                vi = cd['view_intrinsic'][None].repeat((pixel_id.shape[0], 1, 1)) if cd['view_intrinsic'] is not None \
                    else intrinsics
                ve = cd['processed_view_pose'][None].repeat((pixel_id.shape[0], 1, 1))
                _, vrd = geometry.get_rays_tr(pixel_id, vi, ve, mip_nerf_used=False)
                view_dirs = geometry.get_view_dirs_tr(vrd)
            else:
                view_dirs = geometry.get_view_dirs_tr(rays_d)
            ret['view_dirs'] = view_dirs
        o_s, d_s = (rays_o_ndc, rays_d_ndc) if self.ndc else (rays_o, rays_d)
        augment = self.augmentations_needed and self.training and (mode != 'test_camera_params_optimization')

        def composite(net, z):
            if self.ndc:
                return self.volume_rendering(net, z_vals_ndc=z, rays_d_ndc=rays_d_ndc, rays_o=rays_o, rays_d=rays_d)
            return self.volume_rendering(net, z_vals=z, rays_d=rays_d)

        weights_coarse = z_coarse = None
        if self.coarse_model_needed:
            z_coarse = self.get_z_vals_coarse(input_dict)
            # the sample points o + d z are generated inside the MLP kernel (srf_mlp_forward_rays)
            net_in = {'rays_o': o_s, 'rays_d': d_s, 'z_vals': z_coarse}
            if self.configs['model']['coarse_model']['use_view_dirs']:
                net_in['view_dirs'] = view_dirs
            net_out = self.run_network(net_in, self.coarse_model)
            outs = composite(net_out, z_coarse)
            weights_coarse = outs['weights']
            ret['z_vals_coarse'] = z_coarse
            for k in outs:
                ret[f'{k}_coarse'] = outs[k]
            if retraw:
                for k in net_out:
                    ret[f'raw_{k}_coarse'] = net_out[k]
            if augment:
                for aug in self.augmented_models:
                    if aug['coarse_model'] is not None:
                        a_in = dict(net_in)
                        if aug['coarse_model'].mlp_configs['use_view_dirs']:
                            a_in['view_dirs'] = view_dirs
                        a_out = self.run_network(a_in, aug['coarse_model'])
                        a_outs = composite(a_out, z_coarse)
                        for k in a_outs:
                            ret[f"{aug['name']}_{k}_coarse"] = a_outs[k]
                        if retraw:
                            for k in a_out:
                                ret[f"{aug['name']}_raw_{k}_coarse"] = a_out[k]
        if self.fine_model_needed:
            z_fine = self.get_z_vals_fine(z_coarse, weights_coarse)
            net_in = {'rays_o': o_s, 'rays_d': d_s, 'z_vals': z_fine}
            if self.configs['model']['fine_model']['use_view_dirs']:
                net_in['view_dirs'] = view_dirs
            net_out = self.run_network(net_in, self.fine_model)
            outs = composite(net_out, z_fine)
            ret['z_vals_fine'] = z_fine
            for k in outs:
                ret[f'{k}_fine'] = outs[k]
            if retraw:
                for k in net_out:
                    ret[f'raw_{k}_fine'] = net_out[k]
            if augment:
                for aug in self.augmented_models:
                    if aug['fine_model'] is not None:
                        a_out = self.run_network(net_in, aug['fine_model'])
                        a_outs = composite(a_out, z_fine)
                        for k in a_outs:
                            ret[f"{aug['name']}_{k}_fine"] = a_outs[k]
                        if retraw:
                            for k in a_out:
                                ret[f"{aug['name']}_raw_{k}_fine"] = a_out[k]
        if not retraw:
            for stage, needed in (('coarse', self.coarse_model_needed), ('fine', self.fine_model_needed)):
                if needed:
                    for k in ('z_vals', 'alpha', 'visibility', 'weights'):
                        ret.pop(f'{k}_{stage}', None)
            if self.augmentations_needed and self.training:
                for aug in self.augmented_models:
                    for stage in ('coarse', 'fine'):
                        if aug[f'{stage}_model'] is not None:
                            for k in ('alpha', 'visibility', 'weights'):
                                ret.pop(f"{aug['name']}_{k}_{stage}", None)
        return ret

    # --------------------------------------------------------------- sampling
    def _rand(self, shape, device):
        return torch.rand(shape, device=device, generator=self.generator)

    def get_z_vals_coarse(self, input_dict: dict):
        """src/models/SimpleNeRF17.py:330-358 on srf_stratified_samples."""
        num_rays = input_dict['pixel_id'].shape[0]
        dev = input_dict['pixel_id'].device
        if not self.ndc:
            near, far = self.model_configs['near'], self.model_configs['far']
        else:
            near, far = self.model_configs['near_ndc'], self.model_configs['far_ndc']
        perturb = self.training and self.configs['model']['perturb']
        n = self.configs['model']['coarse_model']['num_samples']
        u = self._rand((num_rays, n), dev) if perturb else None
        return sampling.stratified_z_vals(num_rays, n, near, far, lindisp=self.configs['model']['lindisp'], u=u,
                                          device=dev)

    def get_z_vals_fine(self, z_vals_coarse, weights_coarse):
        """src/models/SimpleNeRF17.py:360-371 on srf_fine_z_vals (CDF, searchsorted, merge)."""
        n_fine = self.configs['model']['fine_model']['num_samples']
        perturb = self.configs['model']['perturb'] and self.training
        u = self._rand((z_vals_coarse.shape[0], n_fine), z_vals_coarse.device) if perturb else None
        return sampling.fine_z_vals(z_vals_coarse, weights_coarse, n_fine, u=u)

    @staticmethod
    def sample_pdf(bins, weights, N_samples, det=False):
        """src/models/SimpleNeRF17.py:384-417 on srf_sample_pdf."""
        u = None if det else torch.rand(list(bins.shape[:-1]) + [N_samples], device=bins.device)
        return sampling.sample_pdf(bins.contiguous(), weights.contiguous(), N_samples, u=u)

    # --------------------------------------------------------------- network
    def run_network(self, input_dict, nerf_mlp):
        """src/models/SimpleNeRF17.py:419-448: flatten, one fused MLP launch, reshape.  Besides the
        reference's {'pts': [R,S,3]} it takes the ray form {'rays_o', 'rays_d': [R,3], 'z_vals': [R,S]},
        whose points (src/models/SimpleNeRF17.py:213-215) the kernel generates itself."""
        if 'z_vals' in input_dict:
            lead = list(input_dict['z_vals'].shape)
            net_in = {k: input_dict[k] for k in ('rays_o', 'rays_d', 'z_vals')}
            if nerf_mlp.mlp_configs['use_view_dirs']:
                net_in['view_dirs'] = input_dict['view_dirs']  # one per ray: expanded in the kernel
        else:
            pts = input_dict['pts']
            lead = list(pts.shape[:-1])
            net_in = {'pts': pts.reshape(-1, pts.shape[-1])}
            if nerf_mlp.mlp_configs['use_view_dirs']:
                vd = input_dict['view_dirs']
                if vd.ndim == 2:
                    vd = vd[:, None].expand(pts.shape)
                net_in['view_dirs'] = vd.reshape(-1, vd.shape[-1])
        out = self.batchify(nerf_mlp)(net_in)
        return {k: v.reshape(lead + list(v.shape[1:])) for k, v in out.items()}

    def batchify(self, nerf_mlp):
        """src/models/SimpleNeRF17.py:450-484.  The fused kernel streams any number of points,
        so netchunk slicing is unnecessary; the MLP is called once."""
        return nerf_mlp

    def volume_rendering(self, network_output_dict, z_vals=None, rays_o=None, rays_d=None, z_vals_ndc=None,
                         rays_d_ndc=None, sec_views_vis=False):
        """src/models/SimpleNeRF17.py:486-539 (+ convert_depth_from_ndc :541-558) on srf_composite."""
        sigma = network_output_dict['sigma'][..., 0]
        rgb = network_output_dict['rgb']
        if self.ndc:
            return compositing.composite(sigma, rgb, z_vals_ndc, rays_o, rays_d, rays_d_ndc, ndc=True,
                                         white_bkgd=self.configs['model']['white_bkgd'])
        return compositing.composite(sigma, rgb, z_vals, rays_d, rays_d, None, ndc=False,
                                     white_bkgd=self.configs['model']['white_bkgd'])

    @staticmethod
    def convert_depth_from_ndc(z_vals_ndc, rays_o, rays_d):
        """src/models/SimpleNeRF17.py:541-558 (host formula; the kernels fuse it)."""
        near = 1
        oz, dz = rays_o[..., 2:3], rays_d[..., 2:3]
        tn = -(near + oz) / dz
        constant = torch.where(z_vals_ndc == 1., 1e-3, 0.)
        return (oz + tn * dz) / dz * (1 / (1 - z_vals_ndc + constant) - 1) + tn

    @staticmethod
    def merge_mini_batch_data(data_chunks: dict):
        merged = {}
        for key in data_chunks:
            if isinstance(data_chunks[key][0], torch.Tensor):
                chunks = data_chunks[key]
                merged[key] = chunks[0] if len(chunks) == 1 else torch.cat(chunks, dim=0)
            else:
                raise NotImplementedError
        return merged

    @classmethod
    def deep_dict_copy(cls, input_dict: dict):
        return {k: cls.deep_dict_copy(v) if isinstance(v, dict) else v for k, v in input_dict.items()}


class MLP(torch.nn.Module):
    """Parameters exactly as the reference MLP (src/models/SimpleNeRF17.py:616-667); forward runs
    the fused HIP kernel (srf_mlp_forward) in configs['model']['mlp_precision']."""

    def __init__(self, name, configs, mlp_configs, model_configs):
        super().__init__()
        self.name = name
        self.configs = configs
        self.mlp_configs = mlp_configs
        self.model_configs = model_configs
        self.Dp = mlp_configs['points_net_depth']
        self.Dv = mlp_configs['views_net_depth']
        self.Wp = mlp_configs['points_net_width']
        self.Wv = mlp_configs['views_net_width']
        if (self.Dp, self.Dv, self.Wp, self.Wv) != (8, 1, 256, 128):
            raise NotImplementedError('the fused MLP kernels are specialised to depth 8 x 256 with one 128-wide '
                                      f'view layer; got {(self.Dp, self.Wp, self.Dv, self.Wv)}')
        deg = mlp_configs['points_positional_encoding_degree']
        self.pts_input_dim = 3 * (2 * deg + 1)
        self.views_input_dim = 3 * (2 * mlp_configs['views_positional_encoding_degree'] + 1) \
            if mlp_configs['use_view_dirs'] else 0
        sig_deg = mlp_configs.get('points_sigma_positional_encoding_degree')
        if sig_deg is not None:
            self.pts_input_dim = (2 * sig_deg + 1) * 3
            self.views_input_dim += (2 * deg + 1) * 3 - self.pts_input_dim
        self.skips = [4]
        self.view_dep_rgb = mlp_configs['view_dependent_rgb']
        self.predict_visibility = mlp_configs['predict_visibility']
        if self.predict_visibility:
            raise NotImplementedError('predict_visibility is not supported by the HIP MLP kernels')
        if self.view_dep_rgb != mlp_configs['use_view_dirs']:
            raise NotImplementedError('use_view_dirs and view_dependent_rgb must agree for the fused kernels')
        self.view_dep_outputs = self.view_dep_rgb
        self.raw_noise_std = configs['model']['raw_noise_std']
        self.pts_linears = torch.nn.ModuleList(
            [torch.nn.Linear(self.pts_input_dim, self.Wp)] +
            [torch.nn.Linear(self.Wp, self.Wp) if i not in self.skips else
             torch.nn.Linear(self.Wp + self.pts_input_dim, self.Wp) for i in range(self.Dp - 1)])
        if self.view_dep_outputs:
            self.views_linears = torch.nn.ModuleList([torch.nn.Linear(self.views_input_dim + self.Wp, self.Wv)])
            self.pts_output_linear = torch.nn.Linear(self.Wp, 1)
            self.feature_linear = torch.nn.Linear(self.Wp, self.Wp)
            self.views_output_linear = torch.nn.Linear(self.Wv, 3)
        else:
            self.pts_output_linear = torch.nn.Linear(self.Wp, 4)
        self.spec = hip_mlp.MLPSpec(pts_pe_degree=deg, dirs_pe_degree=mlp_configs.get('views_positional_encoding_degree', 4),
                                    sigma_pe_degree=sig_deg, view_dependent=self.view_dep_rgb,
                                    prec=configs['model'].get('mlp_precision', 'bf16'))

    def get_trainable_parameters(self, optimizer_configs):
        """src/models/SimpleNeRF17.py:684-694."""
        p = torch.nn.ParameterList()
        p.extend(self.parameters())
        return [{'name': f'{self.name}_network_params', 'params': p, 'lr': optimizer_configs['lr_initial']}]

    def layer_params(self):
        lins = list(self.pts_linears) + [getattr(self, 'feature_linear', None), self.pts_output_linear,
                                          self.views_linears[0] if self.view_dep_outputs else None,
                                          getattr(self, 'views_output_linear', None)]
        return [None if l is None else l.weight for l in lins], [None if l is None else l.bias for l in lins]

    def forward(self, input_batch: dict):
        ray_form = 'z_vals' in input_batch  # rays_o, rays_d [R,3], z_vals [R,S]: points made in the kernel
        ref = input_batch['z_vals'] if ray_form else input_batch['pts']
        n = ref.numel() if ray_form else ref.shape[0]
        noise = None
        if self.training and self.raw_noise_std > 0.:
            noise = torch.randn(n, device=ref.device) * self.raw_noise_std
        ws, bs = self.layer_params()
        if ray_form:
            sigma, rgb = hip_mlp.mlp_forward_rays(self.spec, ws, bs, input_batch['rays_o'], input_batch['rays_d'], ref,
                                                  input_batch.get('view_dirs'), noise)
        else:
            sigma, rgb = hip_mlp.mlp_forward(self.spec, ws, bs, ref, input_batch.get('view_dirs'), noise)
        out = {'sigma': sigma[:, None]}
        out['rgb_view_dependent' if self.view_dep_rgb else 'rgb_view_independent'] = rgb
        out['rgb'] = rgb
        return out


class IntrinsicsLearner(torch.nn.Module):
    """Frozen per-frame intrinsics (src/models/SimpleNeRF17.py:788-814)."""

    def __init__(self, initial_intrinsics, learn_focal):
        super().__init__()
        self.name = self.__class__.__name__
        self.initial_intrinsics = torch.nn.Parameter(torch.from_numpy(numpy.asarray(initial_intrinsics).astype(numpy.float32)),
                                                     requires_grad=False)
        if learn_focal:
            raise NotImplementedError

    def forward(self, cam_id):
        return self.initial_intrinsics[cam_id]


class ExtrinsicsLearner(torch.nn.Module):
    """Per-frame extrinsics with an (unlearned here) axis-angle delta (src/models/SimpleNeRF17.py:817-912)."""

    def __init__(self, initial_extrinsics, learn_rotation, learn_translation):
        super().__init__()
        self.name = self.__class__.__name__
        e = numpy.asarray(initial_extrinsics).astype(numpy.float32)
        self.num_frames = e.shape[0]
        self.initial_extrinsics = torch.nn.Parameter(torch.from_numpy(e), requires_grad=False)
        self.r = torch.nn.Parameter(torch.zeros(size=(self.num_frames, 3), dtype=torch.float32), requires_grad=learn_rotation)
        self.t = torch.nn.Parameter(torch.zeros(size=(self.num_frames, 3), dtype=torch.float32),
                                    requires_grad=learn_translation)

    def forward(self, cam_id):
        r, t = self.r[cam_id], self.t[cam_id]
        return self.initial_extrinsics[cam_id] @ self.make_extrinsics(r, t)

    def make_extrinsics(self, r, t):
        R = self.Exp(r)
        c2w = torch.cat([R, t.unsqueeze(2)], dim=2)
        c2w = torch.cat([c2w, torch.zeros_like(c2w[:, 0:1])], dim=1)
        c2w[:, 3, 3] = 1.0
        return torch.linalg.inv_ex(c2w)[0]  # torch.linalg.inv without the host-synchronising check

    @staticmethod
    def Exp(r):
        zero = torch.zeros((r.shape[0], 1), dtype=torch.float32, device=r.device)
        skew = torch.stack([torch.cat([zero, -r[:, 2:3], r[:, 1:2]], 1), torch.cat([r[:, 2:3], zero, -r[:, 0:1]], 1),
                            torch.cat([-r[:, 1:2], r[:, 0:1], zero], 1)], dim=2)
        nr = r.norm(dim=1) + 1e-15
        eye = torch.eye(3, dtype=torch.float32, device=r.device)
        return eye[None] + (torch.sin(nr) / nr)[:, None, None] * skew + \
            ((1 - torch.cos(nr)) / nr ** 2)[:, None, None] * (skew @ skew)


__all__ = ['SimpleNeRFHip', 'MLP', 'IntrinsicsLearner', 'ExtrinsicsLearner', 'SrfError']
