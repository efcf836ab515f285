"""Drop-in replacement for the Simple-TensoRF model on the MI355X HIP kernels.

Selected by the reference's plugin loader (src/models/ModelFactory02.py:10-22) with
``configs['model']['name'] = 'SimpleTensoRFHip09'`` (class ``SimpleTensoRFHip``).  The class
surface, config keys, output-dict keys and state_dict names/shapes are those of SimpleTensoRF
(src/models/SimpleTensoRF09.py:17-1421) for the vector-matrix and CP decompositions.  The hot
path runs on the HIP kernels:

  get_z_vals_coarse (NDC)        -> srf_stratified_samples             (SimpleTensoRF09.py:364-381)
  get_z_vals_coarse (AABB)       -> srf_aabb_samples                   (:382-400)
  get_z_vals_fine                -> srf_fine_z_vals                    (:403-414, sample_pdf :427-460)
  validity + alpha mask + density-> srf_vm_density_fwd / _bwd          (:701-721, :1214-1239, :1342-1345)
                                    (CP: srf_cp_density_fwd / _bwd      :1043-1062)
  get_volume_rendering_weights   -> srf_composite_fwd / _bwd (no rgb)  (:767-811)
  surface mask + colour gather   -> srf_threshold_compact,
                                    srf_vm_appearance_fwd / _bwd       (:726, :1241-1263)
                                    (CP: srf_cp_appearance_fwd / _bwd   :1064-1078)
  volume_render                  -> srf_volume_render_fwd / _bwd       (:814-817)

  basis_matrix_color + color MLP -> srf_color_forward / _backward        (:1262-1270, :1370-1421)

The factor
parameters keep the reference's shapes but are stored texel-major (component axis innermost),
which is the layout the gather kernels read.  Random draws of the path (t_rand :376, the
per-ray offset :396, sample_pdf's u) come from the GPU generator and are passed to the kernels.
The periodic model modifications (alpha-mask update with its 3x3x3 max pool, shrink to the
occupied box, factor upsampling, optimiser reconfiguration; :820-944) run on
srf_dense_grid_points, the density gather, srf_alpha_mask_build and srf_factor_upsample.
Options outside the path raise NotImplementedError: visibility prediction.

``configs['model']['color_precision']`` in {'fp32', 'bf16_split', 'bf16'} (default 'fp32', the
reference's arithmetic) selects the operands of the colour network's MFMA products: split BF16
keeps FP32-level accuracy (three planes forward, two backward), 'bf16' rounds them to BF16.
"""
from __future__ import annotations

import numpy
import torch
import torch.nn.functional as F
from torch.nn import ModuleDict, ModuleList

from srf_hip import color as hip_color, compositing, geometry, sampling, tensorf, tensorf_mods

from .SimpleNeRFHip17 import ExtrinsicsLearner, IntrinsicsLearner

# matrix_axes / vector_axes of the VM decomposition (SimpleTensoRF09.py:1131-1132)
MATRIX_AXES = [[0, 1], [0, 2], [1, 2]]
VECTOR_AXES = [2, 1, 0]


class SimpleTensoRFHip(torch.nn.Module):
    def __init__(self, configs: dict, model_configs: dict):
        """src/models/SimpleTensoRF09.py:17-40."""
        super().__init__()
        self.configs = configs
        self.model_configs = model_configs
        self.ndc = self.configs['data_loader']['ndc']
        self.coarse_model_needed = 'coarse_model' in self.configs['model']
        self.fine_model_needed = 'fine_model' in self.configs['model']
        self.predict_visibility = (self.coarse_model_needed and self.configs['model']['coarse_model']['predict_visibility']) or \
                                  (self.fine_model_needed and self.configs['model']['fine_model']['predict_visibility'])
        if self.predict_visibility:
            raise NotImplementedError('visibility prediction is outside the HIP hot path')
        self.nerf_synthetic = self.configs['database'] == 'NeRF_Synthetic'
        self.coarse_model = None
        self.fine_model = None
        self.intrinsics_learner = None
        self.extrinsics_learner = None
        self.augmentations_needed = 'augmentations' in configs['model']
        if self.augmentations_needed:
            self.augmented_models = []
            self.augmented_models_nn = []
        self.generator = None  # optional torch.Generator on the ray device for the path's draws
        self.build_nerf()
        self.optimizers = None
        self.train_data_preprocessor = None

    def build_nerf(self):
        """src/models/SimpleTensoRF09.py:42-82."""
        if self.coarse_model_needed:
            self.coarse_model = self.get_tensor_model('coarse_model', self.configs, self.configs['model']['coarse_model'],
                                                      self.model_configs)
        if self.fine_model_needed:
            self.fine_model = self.get_tensor_model('fine_model', self.configs, self.configs['model']['fine_model'],
                                                    self.model_configs)
        self.intrinsics_learner = IntrinsicsLearner(numpy.array(self.model_configs['intrinsics']),
                                                    learn_focal=self.configs['model']['learn_camera_focal_length'])
        self.extrinsics_learner = ExtrinsicsLearner(numpy.array(self.model_configs['extrinsics']),
                                                    learn_rotation=self.configs['model']['learn_camera_rotation'],
                                                    learn_translation=self.configs['model']['learn_camera_translation'])
        if self.augmentations_needed:
            for aug_cfg in self.configs['model']['augmentations']:
                aug_name = aug_cfg['name']
                coarse = fine = None
                nn = ModuleDict()
                if 'coarse_model' in aug_cfg:
                    coarse = self.get_tensor_model(f'{aug_name}_coarse_model', self.configs, aug_cfg['coarse_model'],
                                                   self.model_configs)
                    nn['coarse_model'] = coarse
                if 'fine_model' in aug_cfg:
                    fine = self.get_tensor_model(f'{aug_name}_fine_model', self.configs, aug_cfg['fine_model'],
                                                 self.model_configs)
                    nn['fine_model'] = fine
                self.augmented_models.append({'name': aug_name, 'coarse_model': coarse, 'fine_model': fine})
                self.augmented_models_nn.append(nn)
            self.augmented_models_nn = ModuleList(self.augmented_models_nn)

    def get_trainable_parameters(self, optimizer_configs):
        """src/models/SimpleTensoRF09.py:84-97."""
        params = []
        if self.coarse_model_needed:
            params.extend(self.coarse_model.get_trainable_parameters(optimizer_configs))
        if self.fine_model_needed:
            params.extend(self.fine_model.get_trainable_parameters(optimizer_configs))
        if self.augmentations_needed:
            for aug in self.augmented_models:
                for key in ('coarse_model', 'fine_model'):
                    if aug[key] is not None:
                        params.extend(aug[key].get_trainable_parameters(optimizer_configs))
        return params

    def rebuild_camera_params_learners(self, *, intrinsics: numpy.ndarray = None, extrinsics=None, device):
        """src/models/SimpleTensoRF09.py:123-133."""
        if intrinsics is not None:
            self.intrinsics_learner = IntrinsicsLearner(intrinsics, learn_focal=self.configs['model']['learn_camera_focal_length']).to(device)
        if extrinsics is not None:
            self.extrinsics_learner = ExtrinsicsLearner(extrinsics, learn_rotation=self.configs['model']['learn_camera_rotation'],
                                                        learn_translation=self.configs['model']['learn_camera_translation']).to(device)

    def forward(self, input_batch: dict, *, retraw: bool = False, sec_views_vis: bool = False, mode: str = None):
        """src/models/SimpleTensoRF09.py:135-162."""
        input_batch = self.deep_dict_copy(input_batch)
        if 'common_data' in input_batch.keys():
            for key in input_batch['common_data'].keys():
                if isinstance(input_batch['common_data'][key], torch.Tensor):
                    input_batch['common_data'][key] = input_batch['common_data'][key][0]
        if self.training and (mode != 'test_camera_params_optimization') and (input_batch.get('sub_batch_index', 0) == 0):
            self.run_model_modifications(input_batch.get('iter_num'))
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
        input_batch['intrinsics_inv'] = torch.linalg.inv_ex(self.intrinsics_learner(k_cams))[0][image_id]
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
        """src/models/SimpleTensoRF09.py:168-192."""
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
        """src/models/SimpleTensoRF09.py:194-351."""
        ret = {}
        pixel_id = input_dict['pixel_id']
        intrinsics, extrinsics = input_dict['intrinsics'], input_dict['extrinsics']
        num_rays = pixel_id.shape[0]
        rays_o, rays_d = geometry.get_rays_tr(pixel_id, intrinsics, extrinsics, mip_nerf_used=True,
                                              nerf_synthetic=self.nerf_synthetic,
                                              intrinsic_inv=input_dict.get('intrinsics_inv'))
        if torch.is_grad_enabled() and (rays_o.requires_grad or rays_d.requires_grad):
            # a learned camera differentiates the loss through the sample points into the rays; the
            # factor gathers detach their coordinates (src/models/SimpleTensoRF09.py:1231-1235), and
            # the kernels give the rays no gradient
            raise NotImplementedError('camera gradients through the rays are not supported by the HIP kernels')
        rays_o = rays_o.clone()
        rays_d = rays_d.clone()
        rays_o[:, :1] *= -1
        rays_d[:, :1] *= -1
        if self.nerf_synthetic:
            rays_o[:, :1] *= -1
            rays_d[:, :1] *= -1
            rays_d = rays_d / torch.linalg.norm(rays_d, dim=1, keepdim=True)
        ret['rays_o'], ret['rays_d'] = rays_o, rays_d
        rays_o_ndc = rays_d_ndc = None
        if self.ndc:
            rays_o_ndc, rays_d_ndc = geometry.get_ndc_rays_tr(rays_o, rays_d, self.model_configs['resolution'],
                                                              intrinsics, near=self.model_configs['near'])
            ret['rays_o_ndc'], ret['rays_d_ndc'] = rays_o_ndc, rays_d_ndc
        view_dirs = None
        if (self.coarse_model_needed and self.configs['model']['coarse_model']['use_view_dirs']) or \
                (self.fine_model_needed and self.configs['model']['fine_model']['use_view_dirs']):
            if mode == 'static_camera':
                cd = input_dict['common_data']
                vi = cd['view_intrinsic'][None].repeat((num_rays, 1, 1)) if cd['view_intrinsic'] is not None \
                    else intrinsics
                ve = cd['processed_view_pose'][None].repeat((num_rays, 1, 1))
                vro, vrd = geometry.get_rays_tr(pixel_id, vi, ve, mip_nerf_used=False)
                if not self.ndc:
                    view_dirs = geometry.get_view_dirs_tr(vrd)
                else:
                    _, vrd_ndc = geometry.get_ndc_rays_tr(vro, vrd, self.model_configs['resolution'], vi,
                                                          near=self.model_configs['near'])
                    view_dirs = geometry.get_view_dirs_tr(vrd_ndc)
            else:
                view_dirs = geometry.get_view_dirs_tr(rays_d if not self.ndc else rays_d_ndc)
            ret['view_dirs'] = view_dirs
        o_s, d_s = (rays_o_ndc, rays_d_ndc) if self.ndc else (rays_o, rays_d)
        augment = self.augmentations_needed and self.training and (mode != 'test_camera_params_optimization')

        def model_input(z, use_view_dirs):
            d = {'iter_num': input_dict.get('iter_num'), 'pts': o_s[..., None, :] + d_s[..., None, :] * z[..., :, None],
                 'rays_o': rays_o, 'rays_d': rays_d}
            if not self.ndc:
                d['z_vals'] = z
            else:
                d['z_vals_ndc'] = z
                d['rays_d_ndc'] = rays_d_ndc
            if use_view_dirs:
                d['view_dirs'] = view_dirs
            return d

        weights_coarse = z_coarse = None
        if self.coarse_model_needed:
            z_coarse = self.get_z_vals_coarse(input_dict, ret, self.coarse_model)
            m_in = model_input(z_coarse, self.configs['model']['coarse_model']['use_view_dirs'])
            outs = self.coarse_model(m_in, retraw)
            weights_coarse = outs['weights']
            ret['z_vals_coarse'] = z_coarse
            for k in outs:
                ret[f'{k}_coarse'] = outs[k]
            if augment:
                for aug in self.augmented_models:
                    if aug['coarse_model'] is not None:
                        a_outs = aug['coarse_model'](m_in, retraw)
                        for k in a_outs:
                            ret[f"{aug['name']}_{k}_coarse"] = a_outs[k]
        if self.fine_model_needed:
            z_fine = self.get_z_vals_fine(z_coarse, weights_coarse)
            m_in = model_input(z_fine, self.configs['model']['fine_model']['use_view_dirs'])
            outs = self.fine_model(m_in, retraw)
            ret['z_vals_fine'] = z_fine
            for k in outs:
                ret[f'{k}_fine'] = outs[k]
            if augment:
                for aug in self.augmented_models:
                    if aug['fine_model'] is not None:
                        a_outs = aug['fine_model'](m_in, retraw)
                        for k in a_outs:
                            ret[f"{aug['name']}_{k}_fine"] = a_outs[k]
        if not retraw:
            for stage, needed in (('coarse', self.coarse_model_needed), ('fine', self.fine_model_needed)):
                if needed:
                    for k in ('z_vals', 'alpha', 'visibility', 'weights'):
                        ret.pop(f'{k}_{stage}', None)
            if augment:
                for aug in self.augmented_models:
                    for stage in ('coarse', 'fine'):
                        if aug[f'{stage}_model'] is not None:
                            for k in ('alpha', 'visibility', 'weights'):
                                ret.pop(f"{aug['name']}_{k}_{stage}", None)
        return ret

    # ------------------------------------------------------------------ sampling
    def _rand(self, shape, device):
        return torch.rand(shape, device=device, generator=self.generator)

    def get_z_vals_coarse(self, input_dict: dict, output_dict: dict, tensor_model=None):
        """src/models/SimpleTensoRF09.py:353-401."""
        num_rays = input_dict['pixel_id'].shape[0]
        dev = input_dict['pixel_id'].device
        if not self.ndc:
            near, far = self.model_configs['near'], self.model_configs['far']
        else:
            near, far = self.model_configs['near_ndc'], self.model_configs['far_ndc']
        perturb = self.training and self.configs['model']['perturb']
        n = tensor_model.host_params()['num_samples']
        if self.ndc:
            u = self._rand((num_rays, n), dev) if perturb else None
            return sampling.stratified_z_vals(num_rays, n, near, far, lindisp=self.configs['model']['lindisp'], u=u,
                                              device=dev)
        u = self._rand((num_rays,), dev) if perturb else None
        hp = tensor_model.host_params()
        return tensorf.aabb_samples(output_dict['rays_o'], output_dict['rays_d'], hp['bbox'], near, far,
                                    hp['step_size'], n, u=u)

    def get_z_vals_fine(self, z_vals_coarse, weights_coarse):
        """src/models/SimpleTensoRF09.py:403-414 on srf_fine_z_vals."""
        n_fine = self.configs['model']['fine_model']['num_samples']
        perturb = self.configs['model']['perturb'] and self.training
        u = self._rand((z_vals_coarse.shape[0], n_fine), z_vals_coarse.device) if perturb else None
        return sampling.fine_z_vals(z_vals_coarse, weights_coarse, n_fine, u=u)

    def run_model_modifications(self, iter_num):
        """src/models/SimpleTensoRF09.py:462-473."""
        for m in self._tensor_models():
            m.run_model_modifications(iter_num)

    def _tensor_models(self):
        ms = [m for m in (self.coarse_model, self.fine_model) if m is not None]
        if self.augmentations_needed:
            for aug in self.augmented_models:
                ms += [aug[k] for k in ('coarse_model', 'fine_model') if aug[k] is not None]
        return ms

    @staticmethod
    def merge_mini_batch_data(data_chunks: dict):
        merged = {}
        for key in data_chunks:
            if isinstance(data_chunks[key][0], torch.Tensor):
                merged[key] = torch.cat(data_chunks[key], dim=0)
            else:
                raise NotImplementedError
        return merged

    @classmethod
    def deep_dict_copy(cls, input_dict: dict):
        return {k: cls.deep_dict_copy(v) if isinstance(v, dict) else v for k, v in input_dict.items()}

    @staticmethod
    def get_tensor_model(name, configs, tensor_configs, model_configs):
        """src/models/SimpleTensoRF09.py:535-544."""
        kind = tensor_configs['decomposition_type']
        if kind == 'VectorMatrix':
            return VmDecomposedTensor(name, configs, tensor_configs, model_configs)
        if kind == 'CandecompParafac':
            return CpDecomposedTensor(name, configs, tensor_configs, model_configs)
        raise RuntimeError(f'Unknown tensor decomposition: {kind}')


def positional_encoding(x: torch.Tensor, degree: int) -> torch.Tensor:
    """PositionalEncoder with include_input, log sampling and [sin, cos] per frequency
    (SimpleTensoRF09.py:547-579, configured at :1397-1408)."""
    if degree == 0:
        return x
    freqs = 2. ** torch.linspace(0., degree - 1, steps=degree)
    parts = [x]
    for f in freqs:
        parts += [torch.sin(x * f.to(x.device)), torch.cos(x * f.to(x.device))]
    return torch.cat(parts, -1)


class MlpFeaturesColorPredictor(torch.nn.Module):
    """src/models/SimpleTensoRF09.py:1370-1421 (same parameter names: mlp.0, mlp.2, mlp.4)."""

    def __init__(self, tensor_configs, features_dim, features_pe_degree, views_pe_degree, num_units):
        super().__init__()
        self.tensor_configs = tensor_configs
        self.features_pe_degree = features_pe_degree
        self.views_pe_degree = views_pe_degree
        in_dim = features_dim * (2 * features_pe_degree + 1)
        if tensor_configs['use_view_dirs']:
            in_dim += 3 * (2 * views_pe_degree + 1)
        self.input_dim = in_dim
        self.mlp = torch.nn.Sequential(torch.nn.Linear(in_dim, num_units), torch.nn.ReLU(inplace=True),
                                       torch.nn.Linear(num_units, num_units), torch.nn.ReLU(inplace=True),
                                       torch.nn.Linear(num_units, 3), torch.nn.Sigmoid())
        torch.nn.init.constant_(self.mlp[-2].bias, 0)

    def forward(self, input_batch: dict):
        x = positional_encoding(input_batch['features'], self.features_pe_degree)
        if self.tensor_configs['view_dependent_color']:
            x = torch.cat([x, positional_encoding(input_batch['view_dirs'], self.views_pe_degree)], dim=-1)
        return {'color': self.mlp(x)}


class AlphaGridMask(torch.nn.Module):
    """Occupancy grid with the reference's buffers (SimpleTensoRF09.py:1323-1368)."""

    def __init__(self, alpha_volume, bounding_box):
        super().__init__()
        self.register_buffer('alpha_volume', alpha_volume.float().reshape(1, 1, *alpha_volume.shape[-3:]))
        self.register_buffer('bounding_box', bounding_box.float().clone())
        self.register_buffer('bounding_box_size', self.bounding_box[1] - self.bounding_box[0])
        self.register_buffer('resolution', torch.tensor([alpha_volume.shape[-1], alpha_volume.shape[-2],
                                                         alpha_volume.shape[-3]], dtype=torch.long))
        self._register_state_dict_hook(self._save_hook)
        self._register_load_state_dict_pre_hook(self._load_hook)
        self._mask = None

    @staticmethod
    def _save_hook(module, state, prefix, *args, **kwargs):
        state[f'{prefix}alpha_volume'] = state[f'{prefix}alpha_volume'].bool()

    def _load_hook(self, state, prefix, *args, **kwargs):
        state[f'{prefix}alpha_volume'] = state[f'{prefix}alpha_volume'].float()
        self._mask = None

    def kernel_mask(self) -> tensorf.AlphaMask:
        if self._mask is None or self._mask.volume.data_ptr() != self.alpha_volume.data_ptr():
            self._mask = tensorf.AlphaMask(self.alpha_volume, self.bounding_box.flatten().tolist())
        return self._mask

    def sample_alpha(self, pts):
        """src/models/SimpleTensoRF09.py:1342-1345 on srf_alpha_mask_sample."""
        return tensorf.alpha_mask_sample(self.kernel_mask(), pts)


class LowRankTensor(torch.nn.Module):
    """LowRankTensor (src/models/SimpleTensoRF09.py:582-961): geometry buffers, alpha mask, colour
    network and the forward on the HIP kernels.  The decompositions below supply the factor
    parameters and their density / appearance gathers."""

    FACTOR_LISTS = ()  # names of the ParameterLists holding texel-major factors

    def __init__(self, name, configs, tensor_configs, model_configs):
        super().__init__()
        self.name = name
        self.configs = configs
        self.tensor_configs = tensor_configs
        self.model_configs = model_configs
        self.ndc = self.configs['data_loader']['ndc']
        self.predict_visibility = self.tensor_configs['predict_visibility']
        if self.predict_visibility:
            raise NotImplementedError('visibility prediction is outside the HIP hot path')
        counts = list(self.tensor_configs['num_components_density']) + list(self.tensor_configs['num_components_color'])
        if any(int(c) % 4 for c in counts):
            # the reference's configurations use 4, 12, 16 and 48 components per factor
            raise NotImplementedError(f'the factor gathers read components in groups of 4: {counts} are not all '
                                      'multiples of 4')
        self.register_buffer('bounding_box', torch.zeros(size=(2, 3)))
        self.register_buffer('resolution', torch.zeros(size=(3,)))
        self.register_buffer('num_samples', torch.tensor(0))
        self.register_buffer('bounding_box_size', torch.zeros(size=(3,)))
        self.register_buffer('voxel_length', torch.zeros(size=(3,)))
        self.register_buffer('step_size', torch.tensor(0))
        bbox = self.tensor_configs.get('bounding_box', self.model_configs.get('bounding_box'))
        self.bounding_box = torch.tensor(bbox)
        self.resolution = self.compute_resolution_in_voxels(tensor_configs['num_voxels_initial'], self.bounding_box)
        self.update_tensor_params(self.resolution, self.bounding_box)
        self.alpha_mask = None
        self.optimizers = None
        self._host = None
        self._register_load_state_dict_pre_hook(self._load_hook)
        self.build_tensor()
        self.density_kind = tensorf.predictor_code(self.tensor_configs['density_predictor'])
        if self.tensor_configs['color_predictor'] != 'MLP_Features':
            raise NotImplementedError(self.tensor_configs['color_predictor'])
        if bool(self.tensor_configs['use_view_dirs']) != bool(self.tensor_configs['view_dependent_color']):
            raise NotImplementedError('colour MLP with use_view_dirs != view_dependent_color')
        self.color_predictor = MlpFeaturesColorPredictor(
            self.tensor_configs, self.tensor_configs['features_dimension_color'],
            self.tensor_configs['features_positional_encoding_degree'],
            self.tensor_configs['views_positional_encoding_degree'], self.tensor_configs['num_units_color_predictor'])
        tc = self.tensor_configs
        self.color_spec = hip_color.ColorSpec(
            n_features=sum(tc['num_components_color']), basis_dim=tc['features_dimension_color'],
            features_pe=tc['features_positional_encoding_degree'], use_view_dirs=bool(tc['view_dependent_color']),
            views_pe=tc['views_positional_encoding_degree'], hidden=tc['num_units_color_predictor'],
            prec=self.configs['model'].get('color_precision', 'fp32'))
        if not self.color_spec.supported():
            raise NotImplementedError(f'colour network {self.color_spec} is outside the HIP kernels')

    # ------------------------------------------------------------- geometry
    @staticmethod
    def compute_resolution_in_voxels(num_voxels: int, bounding_box: torch.Tensor) -> torch.Tensor:
        """src/models/SimpleTensoRF09.py:626-638."""
        xyz_min, xyz_max = bounding_box
        voxel_size = ((xyz_max - xyz_min).prod() / num_voxels).pow(1 / len(xyz_min))
        return ((xyz_max - xyz_min) / voxel_size).long()

    def compute_num_samples(self, resolution: torch.Tensor, num_voxels_per_sample: float):
        """src/models/SimpleTensoRF09.py:640-650."""
        n = (torch.linalg.norm(resolution.float()) / num_voxels_per_sample).round().long()
        return min(self.tensor_configs['num_samples_max'], n)

    def update_tensor_params(self, new_resolution, new_bounding_box):
        """src/models/SimpleTensoRF09.py:652-665."""
        self.bounding_box = new_bounding_box.float()
        self.bounding_box_size = self.bounding_box[1] - self.bounding_box[0]
        self.resolution = new_resolution.long()
        self.voxel_length = self.bounding_box_size / (self.resolution - 1)
        self.step_size = torch.mean(self.voxel_length) * self.tensor_configs['num_voxels_per_sample']
        self.num_samples = torch.as_tensor(self.compute_num_samples(self.resolution,
                                                                    self.tensor_configs['num_voxels_per_sample']))
        self._host = None

    def _load_hook(self, state, prefix, *args, **kwargs):
        self._host = None
        key = f'{prefix}alpha_mask.alpha_volume'
        if key in state and self.alpha_mask is None:
            vol = state[key].float()
            self.alpha_mask = AlphaGridMask(vol.reshape(vol.shape[-3:]), state[f'{prefix}alpha_mask.bounding_box'])

    def host_params(self) -> dict:
        """bbox, step size and sample count as host values, read once per change of the buffers
        (the reference reads num_samples with .item() on every call, :363)."""
        key = (self.bounding_box._version, self.num_samples._version, self.step_size._version,
               self.bounding_box.data_ptr())
        if self._host is None or self._host[0] != key:
            self._host = (key, {'bbox': self.bounding_box.flatten().tolist(), 'num_samples': int(self.num_samples),
                                'step_size': float(self.step_size)})
        return self._host[1]

    # ------------------------------------------------------------- factors
    def build_tensor(self):
        raise NotImplementedError

    def _density(self, pts, bbox, alpha, offset):
        """sigma [N, S] of every sample (0 where invalid)."""
        raise NotImplementedError

    def _appearance(self, pts, bbox, idx, n_rows, count=None):
        """Appearance features [n_rows, n_features] of the samples idx[:n_rows] (idx[:count] with a
        device-side count, n_rows then being the capacity)."""
        raise NotImplementedError

    def _apply(self, fn, recurse=True):
        """Module.to() / .cuda() rebuild parameters with fn, which makes them contiguous in the
        reference's [1, c, H, W] layout; restore the texel-major storage afterwards."""
        out = super()._apply(fn, recurse)
        for name in self.FACTOR_LISTS:
            plist = getattr(self, name, None)
            if plist is None:
                continue
            for i, p in enumerate(plist):
                if not p.permute(0, 2, 3, 1).is_contiguous():
                    _, c, h, w = p.shape
                    np_ = tensorf.texel_major_param(c, h, w, p.detach(), device=p.device)
                    np_.requires_grad_(p.requires_grad)
                    plist[i] = np_
        self._host = None
        return out

    def run_model_modifications(self, iter_num):
        """src/models/SimpleTensoRF09.py:820-829: the alpha-mask update (and, at its first iteration, the
        shrink of the grid to the occupied box), and the upsampling of the factors with the optimiser
        reconfigured for the new parameters."""
        if not self.training or iter_num is None:
            return
        if iter_num in self.tensor_configs['alpha_mask_update_iters']:
            bounding_box_new = self.update_alpha_mask(iter_num)
            if iter_num == self.tensor_configs['alpha_mask_update_iters'][0]:
                self.shrink_tensor(bounding_box_new)
        if iter_num in self.tensor_configs['tensor_upsampling_iters']:
            self.upsample_model_resolution(iter_num)
            self.reconfigure_optimizer()

    def upsample_model_resolution(self, iter_num):
        """src/models/SimpleTensoRF09.py:831-835 (the subclasses resample their factors)."""
        num_voxels_new = self.get_new_num_voxels(iter_num)
        self.resolution = self.compute_resolution_in_voxels(num_voxels_new, self.bounding_box.cpu()).to(
            self.bounding_box.device)
        self.update_tensor_params(self.resolution, self.bounding_box)

    def get_new_num_voxels(self, iter_num):
        """src/models/SimpleTensoRF09.py:837-847: the voxel count interpolated geometrically between the
        initial and final counts over the upsampling iterations."""
        iters = self.tensor_configs['tensor_upsampling_iters']
        if iter_num not in iters:
            raise RuntimeError('LowRankTensor.get_new_num_voxels() called at invalid iteration number')
        k = iters.index(iter_num) + 1
        lo = numpy.log(self.tensor_configs['num_voxels_initial'])
        hi = numpy.log(self.tensor_configs['num_voxels_final'])
        return int(numpy.round(numpy.exp(lo + (hi - lo) * k / len(iters))))

    def update_alpha_mask(self, iter_num):
        """src/models/SimpleTensoRF09.py:849-870 on the kernels: the densities of the dense grid
        (srf_dense_grid_points, then this decomposition's density gather through the current mask), the
        alpha, transpose, 3x3x3 max pool and threshold in srf_alpha_mask_build; returns the box of the
        occupied voxels."""
        grid = tuple(int(v) for v in self.resolution.tolist())
        hp = self.host_params()
        dev = self.bounding_box.device
        pts = tensorf_mods.dense_grid_points(grid, hp['bbox'], dev)
        mask = self.alpha_mask.kernel_mask() if self.alpha_mask is not None else None
        with torch.no_grad():
            sigma = self._density(pts, hp['bbox'], mask, float(self.tensor_configs.get('density_offset', 0.0)))
        alpha, bounds = tensorf_mods.alpha_mask_build(sigma, grid, hp['step_size'],
                                                      self.tensor_configs['alpha_mask_threshold'])
        self.alpha_mask = AlphaGridMask(alpha, self.bounding_box)
        b = bounds.tolist()
        if b[0] > b[3]:
            raise RuntimeError(f'{self.name}: the alpha-mask update found no occupied voxel')
        gx, gy, gz = grid
        lo = pts[(b[0] * gy + b[1]) * gz + b[2]]
        hi = pts[(b[3] * gy + b[4]) * gz + b[5]]
        return torch.stack((lo, hi))

    def shrink_tensor(self, bounding_box_new):
        """src/models/SimpleTensoRF09.py:899-915: the voxel range [t_l, b_r) that keeps the new box on each
        axis (the box snapped to voxel positions when the mask's resolution differs), the new geometry;
        the subclasses crop their factors.  Returns (t_l, b_r) as host lists."""
        bbox, res = self.bounding_box, self.resolution
        xyz_min, xyz_max = bounding_box_new
        t_l, b_r = (xyz_min - bbox[0]) / self.voxel_length, (xyz_max - bbox[0]) / self.voxel_length
        t_l, b_r = torch.round(torch.round(t_l)).long(), torch.round(b_r).long() + 1
        b_r = torch.stack([b_r, res]).amin(0)
        if not torch.equal(self.alpha_mask.resolution.to(res.device), res):
            t_l_r, b_r_r = t_l / (res - 1), (b_r - 1) / (res - 1)
            corrected = torch.zeros_like(bounding_box_new)
            corrected[0] = (1 - t_l_r) * bbox[0] + t_l_r * bbox[1]
            corrected[1] = (1 - b_r_r) * bbox[0] + b_r_r * bbox[1]
            bounding_box_new = corrected
        self.update_tensor_params(b_r - t_l, bounding_box_new)
        return t_l.tolist(), b_r.tolist()

    def reconfigure_optimizer(self):
        """src/models/SimpleTensoRF09.py:917-944: replace this model's parameter groups in the main
        optimiser by groups of the new (resampled) parameters, dropping the old parameters' state."""
        if not self.optimizers:
            return
        optimizer = self.optimizers['optimizer_nerf']
        opt_cfg = next(c for c in self.configs['optimizers'] if c['name'] == 'optimizer_main')
        groups = self.get_trainable_parameters(opt_cfg)
        names = {g['name'] for g in groups}
        for old in [g for g in optimizer.param_groups if g.get('name') in names]:
            for q in old['params']:
                optimizer.state.pop(q, None)
            optimizer.param_groups.remove(old)
        for g in groups:
            optimizer.add_param_group(g)

    def normalize_points(self, pts: torch.Tensor):
        """src/models/SimpleTensoRF09.py:763-765."""
        return ((pts - self.bounding_box[0]) / self.bounding_box_size) * 2 - 1

    # ------------------------------------------------------------- forward
    def forward(self, input_dict: dict, retraw: bool):
        """LowRankTensor.forward (src/models/SimpleTensoRF09.py:701-761) on the HIP kernels."""
        hp = self.host_params()
        pts = input_dict['pts']
        shape = pts.shape[:-1]
        n_rays, n_samples = shape
        alpha = self.alpha_mask.kernel_mask() if self.alpha_mask is not None else None
        sigma = self._density(pts, hp['bbox'], alpha, float(self.tensor_configs.get('density_offset', 0.0)))
        out = self.get_volume_rendering_weights(input_dict, sigma)
        weights = out['weights']
        vd, per_ray = None, n_samples
        if self.tensor_configs['view_dependent_color']:
            vd = input_dict['view_dirs']
            if vd.ndim != 2:          # per-sample directions [N, S, 3]
                vd, per_ray = vd.reshape(-1, 3), 1
            vd = vd.contiguous()
        rgb_flat = self._surface_colours(pts, hp['bbox'], weights, vd, per_ray)
        rgb = rgb_flat.view(n_rays, n_samples, 3)
        rgb_map = self._background(compositing.volume_render(rgb, weights), out['acc'])
        out['rgb'] = rgb_map
        if retraw:
            out['raw_sigma'] = sigma[..., None]
            out['raw_rgb'] = rgb
        return out

    def forward_ragged(self, input_dict: dict, retraw: bool = False):
        """LowRankTensor.forward (src/models/SimpleTensoRF09.py:701-761, non-NDC branch of
        :767-819) over a ragged batch of occupancy-pruned samples: ray r owns the flat samples
        [ray_offsets[r], ray_offsets[r+1]).  input_dict: pts [P,3], z_vals [P] (metric depths),
        ray_offsets [N+1] int32, rays_o / rays_d / view_dirs [N,3], max_samples (largest count, a
        host int).  Same arithmetic per ray as forward() on a dense batch of that ray alone; the
        samples of a ray are composited with the infinite last interval of :776."""
        hp = self.host_params()
        pts, z, off = input_dict['pts'], input_dict['z_vals'], input_dict['ray_offsets']
        n_pts = pts.shape[0]
        n_rays = off.shape[0] - 1
        alpha = self.alpha_mask.kernel_mask() if self.alpha_mask is not None else None
        sigma = self._density(pts, hp['bbox'], alpha, float(self.tensor_configs.get('density_offset', 0.0)))
        out = compositing.composite(sigma, None, z, input_dict['rays_o'], input_dict['rays_d'], None, ndc=False,
                                    distance_scale=float(self.tensor_configs['distance_scale']), ray_offsets=off,
                                    max_samples=int(input_dict['max_samples']))
        out.pop('rgb')
        weights = out['weights']
        vd = None
        if self.tensor_configs['view_dependent_color']:
            counts = (off[1:] - off[:-1]).long()
            ray_of = torch.repeat_interleave(torch.arange(n_rays, device=pts.device), counts, output_size=n_pts)
            vd = input_dict['view_dirs'][ray_of].contiguous()
        rgb_flat = self._surface_colours(pts, hp['bbox'], weights, vd, 1)
        rgb_map = self._background(compositing.volume_render(rgb_flat, weights, ray_offsets=off), out['acc'])
        out['rgb'] = rgb_map
        if retraw:
            out['raw_sigma'] = sigma[..., None]
            out['raw_rgb'] = rgb_flat
        return out

    def _surface_colours(self, pts, bbox, weights, view_dirs, n_samples):
        """rgb of every sample [P, 3]: the colour network's output on the samples whose weight passes
        ray_marching_weight_threshold, 0 elsewhere (src/models/SimpleTensoRF09.py:726, :1241-1272).
        The surface list and its count stay on the device: the appearance gather and the colour
        network run over the capacity P and stop at the count, so nothing here synchronises the
        host (the reference's pts[validity_mask] and validity_mask.any() at :1247-1250 do) and the
        step can be graph-captured."""
        n_all = weights.numel()
        idx, count = tensorf.threshold_compact(weights, self.tensor_configs['ray_marching_weight_threshold'])
        app = self._appearance(pts, bbox, idx, n_all, count)
        mlp = self.color_predictor.mlp
        return hip_color.color_forward(self.color_spec, app, self.basis_matrix_color.weight,
                                       (mlp[0].weight, mlp[0].bias, mlp[2].weight, mlp[2].bias, mlp[4].weight,
                                        mlp[4].bias), view_dirs=view_dirs, idx=idx, n_samples=n_samples, count=count,
                                       dense_rows=n_all)

    def _background(self, rgb_map, acc):
        """White background (src/models/SimpleTensoRF09.py:746-747): always with white_bkgd, else on a coin
        flip per training call.  The reference flips the coin on the host (torch.rand((1,)) < 0.5); here it is drawn on the
        device and applied arithmetically (rgb + coin (1 - acc), exact for coin 0 and 1), so that the
        step has no host-side branch and replays correctly from a captured graph."""
        if self.configs['model']['white_bkgd']:
            return rgb_map + (1 - acc[..., None])
        if self.training:
            coin = (torch.rand((1,), device=rgb_map.device) < 0.5).to(rgb_map.dtype)
            return rgb_map + coin * (1 - acc[..., None])
        return rgb_map

    def get_volume_rendering_weights(self, input_dict: dict, sigma: torch.Tensor):
        """src/models/SimpleTensoRF09.py:767-811 on srf_composite_fwd (colours composited later)."""
        ds = float(self.tensor_configs['distance_scale'])
        if not self.ndc:
            res = compositing.composite(sigma, None, input_dict['z_vals'], input_dict['rays_o'], input_dict['rays_d'],
                                        None, ndc=False, distance_scale=ds)
        else:
            res = compositing.composite(sigma, None, input_dict['z_vals_ndc'], input_dict['rays_o'],
                                        input_dict['rays_d'], input_dict['rays_d_ndc'], ndc=True, distance_scale=ds)
        res.pop('rgb')
        return res

    @staticmethod
    def volume_render(raw_value, weights):
        """src/models/SimpleTensoRF09.py:814-817."""
        return compositing.volume_render(raw_value, weights)


class VmDecomposedTensor(LowRankTensor):
    """VmDecomposedTensor (src/models/SimpleTensoRF09.py:1126-1272): three planes and three lines
    per field, gathered by srf_vm_density_* / srf_vm_appearance_*."""

    FACTOR_LISTS = ('matrices_density', 'vectors_density', 'matrices_color', 'vectors_color')

    def build_tensor(self):
        """src/models/SimpleTensoRF09.py:1144-1152."""
        self.matrices_density, self.vectors_density = self.create_decomposed_tensor(
            self.tensor_configs['num_components_density'], self.resolution, init_params_scale=0.1)
        self.matrices_color, self.vectors_color = self.create_decomposed_tensor(
            self.tensor_configs['num_components_color'], self.resolution, init_params_scale=0.1)
        self.basis_matrix_color = torch.nn.Linear(sum(self.tensor_configs['num_components_color']),
                                                  self.tensor_configs['features_dimension_color'], bias=False)

    @staticmethod
    def create_decomposed_tensor(num_components, resolution, init_params_scale):
        """src/models/SimpleTensoRF09.py:1154-1165, with texel-major storage."""
        matrices, vectors = [], []
        for i in range(3):
            a0, a1 = MATRIX_AXES[i]
            v = VECTOR_AXES[i]
            h, w, l = int(resolution[a1]), int(resolution[a0]), int(resolution[v])
            c = num_components[i]
            matrices.append(tensorf.texel_major_param(c, h, w, init_params_scale * torch.randn((1, c, h, w))))
            vectors.append(tensorf.texel_major_param(c, l, 1, init_params_scale * torch.randn((1, c, l, 1))))
        return torch.nn.ParameterList(matrices), torch.nn.ParameterList(vectors)

    def get_trainable_parameters(self, optimizer_configs):
        """src/models/SimpleTensoRF09.py:1167-1194."""
        tensor_params = torch.nn.ParameterList()
        network_params = torch.nn.ParameterList()
        tensor_params.extend(self.vectors_density)
        tensor_params.extend(self.matrices_density)
        tensor_params.extend(self.vectors_color)
        tensor_params.extend(self.matrices_color)
        network_params.extend(self.basis_matrix_color.parameters())
        network_params.extend(self.color_predictor.parameters())
        return [
            {'name': f'{self.name}_tensor_params', 'params': tensor_params, 'lr': optimizer_configs['lr_initial_tensor']},
            {'name': f'{self.name}_network_params', 'params': network_params,
             'lr': optimizer_configs['lr_initial_network']},
        ]

    def upsample_model_resolution(self, iter_num):
        """src/models/SimpleTensoRF09.py:1274-1279: new geometry, then every plane and line resampled
        (srf_factor_upsample)."""
        super().upsample_model_resolution(iter_num)
        res = self.resolution.tolist()
        for kind in ('density', 'color'):
            planes, lines = getattr(self, f'matrices_{kind}'), getattr(self, f'vectors_{kind}')
            for i in range(3):
                a0, a1 = MATRIX_AXES[i]
                planes[i] = tensorf_mods.upsample_factor(planes[i], res[a1], res[a0])
                lines[i] = tensorf_mods.upsample_factor(lines[i], res[VECTOR_AXES[i]], 1)

    def shrink_tensor(self, bounding_box_new):
        """src/models/SimpleTensoRF09.py:1304-1321: crop every plane and line to the kept voxel range."""
        t_l, b_r = super().shrink_tensor(bounding_box_new)
        for kind in ('density', 'color'):
            planes, lines = getattr(self, f'matrices_{kind}'), getattr(self, f'vectors_{kind}')
            for i in range(3):
                v = VECTOR_AXES[i]
                lines[i] = tensorf_mods.crop_factor(lines[i], slice(t_l[v], b_r[v]), slice(None))
                a0, a1 = MATRIX_AXES[i]
                planes[i] = tensorf_mods.crop_factor(planes[i], slice(t_l[a1], b_r[a1]), slice(t_l[a0], b_r[a0]))
        return t_l, b_r

    def _density(self, pts, bbox, alpha, offset):
        sigma, _ = tensorf.vm_density(pts, list(self.matrices_density), list(self.vectors_density), bbox,
                                      alpha=alpha, predictor=self.density_kind, offset=offset)
        return sigma

    def _appearance(self, pts, bbox, idx, n_rows, count=None):
        return tensorf.vm_appearance(pts, list(self.matrices_color), list(self.vectors_color), bbox, idx=idx,
                                     n_rows=n_rows, count=count)


class CpDecomposedTensor(LowRankTensor):
    """CpDecomposedTensor (src/models/SimpleTensoRF09.py:964-1127): three lines of one rank per
    field (num_components_*[0]), gathered by srf_cp_density_* / srf_cp_appearance_*."""

    FACTOR_LISTS = ('vectors_density', 'vectors_color')

    def build_tensor(self):
        """src/models/SimpleTensoRF09.py:979-987."""
        self.vectors_density = self.create_decomposed_tensor(self.tensor_configs['num_components_density'],
                                                             self.resolution, init_params_scale=0.1)
        self.vectors_color = self.create_decomposed_tensor(self.tensor_configs['num_components_color'],
                                                           self.resolution, init_params_scale=0.1)
        self.basis_matrix_color = torch.nn.Linear(sum(self.tensor_configs['num_components_color']),
                                                  self.tensor_configs['features_dimension_color'], bias=False)

    @staticmethod
    def create_decomposed_tensor(num_components, resolution, init_params_scale):
        """src/models/SimpleTensoRF09.py:989-996 (every line has num_components[0] components),
        with texel-major storage."""
        vectors = []
        c = num_components[0]
        for i in range(3):
            l = int(resolution[VECTOR_AXES[i]])
            vectors.append(tensorf.texel_major_param(c, l, 1, init_params_scale * torch.randn((1, c, l, 1))))
        return torch.nn.ParameterList(vectors)

    def get_trainable_parameters(self, optimizer_configs):
        """src/models/SimpleTensoRF09.py:998-1023."""
        tensor_params = torch.nn.ParameterList()
        network_params = torch.nn.ParameterList()
        tensor_params.extend(self.vectors_density)
        tensor_params.extend(self.vectors_color)
        network_params.extend(self.basis_matrix_color.parameters())
        network_params.extend(self.color_predictor.parameters())
        return [
            {'name': f'{self.name}_tensor_params', 'params': tensor_params, 'lr': optimizer_configs['lr_initial_tensor']},
            {'name': f'{self.name}_network_params', 'params': network_params,
             'lr': optimizer_configs['lr_initial_network']},
        ]

    def upsample_model_resolution(self, iter_num):
        """src/models/SimpleTensoRF09.py:1093-1098: new geometry, then every line resampled."""
        super().upsample_model_resolution(iter_num)
        res = self.resolution.tolist()
        for kind in ('density', 'color'):
            lines = getattr(self, f'vectors_{kind}')
            for i in range(3):
                lines[i] = tensorf_mods.upsample_factor(lines[i], res[VECTOR_AXES[i]], 1)

    def shrink_tensor(self, bounding_box_new):
        """src/models/SimpleTensoRF09.py:1112-1123: crop every line to the kept voxel range."""
        t_l, b_r = super().shrink_tensor(bounding_box_new)
        for kind in ('density', 'color'):
            lines = getattr(self, f'vectors_{kind}')
            for i in range(3):
                v = VECTOR_AXES[i]
                lines[i] = tensorf_mods.crop_factor(lines[i], slice(t_l[v], b_r[v]), slice(None))
        return t_l, b_r

    def _density(self, pts, bbox, alpha, offset):
        sigma, _ = tensorf.cp_density(pts, list(self.vectors_density), bbox, alpha=alpha, predictor=self.density_kind,
                                      offset=offset)
        return sigma

    def _appearance(self, pts, bbox, idx, n_rows, count=None):
        return tensorf.cp_appearance(pts, list(self.vectors_color), bbox, idx=idx, n_rows=n_rows, count=count)
