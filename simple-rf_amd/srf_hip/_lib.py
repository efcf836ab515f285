"""ctypes binding of libsrf_hip.so (the C-ABI declared in include/srf_hip.h).

The library is loaded from this package directory.  There is no fallback: if the
library is missing or was built for another ABI version, every entry point raises.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import torch

LIB_PATH = Path(__file__).resolve().parent / 'libsrf_hip.so'
ABI_VERSION = 3

_P = C.c_void_p
_I = C.c_int
_L = C.c_int64
_F = C.c_float


class VmFactors(C.Structure):
    """struct srf_vm_factors."""
    _fields_ = [('n_comp', C.c_int * 3), ('plane_h', C.c_int * 3), ('plane_w', C.c_int * 3),
                ('line_len', C.c_int * 3), ('planes', C.c_void_p * 3), ('lines', C.c_void_p * 3)]


class CpFactors(C.Structure):
    """struct srf_cp_factors."""
    _fields_ = [('n_comp', C.c_int), ('line_len', C.c_int * 3), ('lines', C.c_void_p * 3)]


class AlphaGrid(C.Structure):
    """struct srf_alpha_grid."""
    _fields_ = [('volume', C.c_void_p), ('depth', C.c_int), ('height', C.c_int), ('width', C.c_int),
                ('bbox', C.c_float * 6)]

# name -> argtypes (restype is int unless listed in _RESTYPES)
_PROTOS = {
    'srf_abi_version': [],
    'srf_status_string': [_I],
    'srf_selftest_mfma': [_I, _P, _P, _P, _P],
    'srf_selftest_lds_dma': [_P, _P, _P],
    'srf_tune_set': [_I, _L],
    'srf_tune_get': [_I],
    'srf_rays': [_I, _P, _I, _P, _P, _I, _I, _I, _I, _F, _P, _P, _P, _P, _P, _P],
    'srf_stratified_samples': [_I, _I, _F, _F, _I, _P, _P, _P],
    'srf_sample_pdf': [_I, _I, _I, _P, _P, _P, _P, _P, _P, _P, _P],
    'srf_fine_z_vals': [_I, _I, _I, _P, _P, _P, _P, _P, _P],
    'srf_composite_fwd': [_I, _I, _P, _P, _P, _P, _P, _P, _P, _I, _I, _F,
                          _P, _P, _P, _P, _P, _P, _P, _P, _P, _P],
    'srf_composite_bwd': [_I, _I, _P, _P, _P, _P, _P, _P, _P, _I, _I, _F,
                          _P, _P, _P,
                          _P, _P, _P, _P, _P, _P, _P, _P, _P,
                          _P, _P, _P],
    'srf_mlp_packed_bytes': [_P],
    'srf_mlp_pack': [_P, _P, _P, _P, _P],
    'srf_mlp_workspace_bytes': [_P, _I],
    'srf_mlp_forward': [_P, _P, _I, _P, _P, _P, _P, _P, _P, _P],
    'srf_mlp_forward_rays': [_P, _P, _I, _I, _P, _P, _P, _P, _P, _P, _P, _P, _P],
    'srf_mlp_backward': [_P, _P, _I, _P, _P, _P, _P, _P, _P, _P, _P],
    'srf_mlp_backward_data': [_P, _P, _I, _P, _P, _P, _P, _P, _P],
    'srf_mlp_weight_grads': [_P, _I, _P, _P, _P, _P],
    'srf_vm_density_fwd': [_P, _P, _P, _L, _P, _I, _F, _P, _P, _P, _P],
    'srf_vm_density_bwd': [_P, _P, _P, _P, _L, _P, _P, _P, _I, _F, _P, _P, _P],
    'srf_vm_appearance_fwd': [_P, _P, _L, _P, _P, _P, _P, _P],
    'srf_vm_appearance_bwd': [_P, _P, _P, _P, _L, _P, _P, _P, _P, _P, _P],
    'srf_vm_bwd_workspace_bytes': [_P, _L],
    'srf_cp_density_fwd': [_P, _P, _P, _L, _P, _I, _F, _P, _P, _P, _P],
    'srf_cp_density_bwd': [_P, _P, _P, _L, _P, _P, _P, _I, _F, _P, _P],
    'srf_cp_appearance_fwd': [_P, _P, _L, _P, _P, _P, _P, _P],
    'srf_cp_appearance_bwd': [_P, _P, _P, _L, _P, _P, _P, _P, _P],
    'srf_alpha_mask_sample': [_P, _L, _P, _P, _P],
    'srf_compact_workspace_bytes': [_L],
    'srf_threshold_compact': [_L, _P, _F, _P, _P, _P, _P],
    'srf_threshold_compact_noalloc': [_L, _P, _F, _P, _P, _P],
    'srf_adam_step': [_I, _P, _P, _P, _P, _P, _P, _P, _P, _F, _F, _F, _P],
    'srf_volume_render_fwd': [_I, _I, _P, _P, _P, _I, _P, _P],
    'srf_volume_render_bwd': [_I, _I, _P, _P, _P, _I, _P, _P, _P, _P],
    'srf_aabb_samples': [_I, _I, _P, _P, _P, _F, _F, _F, _P, _P, _P],
    'srf_color_packed_bytes': [_P],
    'srf_color_workspace_bytes': [_P, _I],
    'srf_color_pack': [_P, _P, _P, _P, _P, _P, _P, _P, _P, _P],
    'srf_color_forward': [_P, _P, _I, _P, _P, _P, _I, _P, _P, _I, _P, _P],
    'srf_color_backward': [_P, _P, _I, _P, _P, _I, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P],
    'srf_mse_terms_fwd': [_I, _P, _P, _P, _P, _I, _I, _P, _P, _P, _P, _P],
    'srf_mse_terms_bwd': [_I, _P, _P, _P, _P, _I, _I, _P, _P, _P, _P],
    'srf_depth_consistency_fwd': [_I, _P, _P, _P, _P, _P, _P, _I, _P, _P, _P, _I, _I, _I, _I, _I, _F, _I,
                                  _P, _P, _P, _P, _P, _P],
    'srf_depth_consistency_bwd': [_I, _P, _P, _P, _P, _P, _P, _P, _P],
    'srf_dense_grid_points': [_I, _I, _I, _P, _P, _P],
    'srf_alpha_mask_build': [_I, _I, _I, _P, _F, _F, _P, _P, _P],
    'srf_factor_upsample': [_I, _I, _I, _P, _I, _I, _P, _P],
    'srf_tv_workspace_bytes': [_I],
    'srf_tv_fwd': [_I, _P, _P, _P, _P, _F, _P, _P, _P],
    'srf_tv_bwd': [_I, _P, _P, _P, _P, _F, _P, _P, _P],
    'srf_entropy_fwd': [_I, _I, _P, _P, _I, _P, _P, _P, _P, _P, _P],
    'srf_entropy_bwd': [_I, _I, _P, _I, _P, _P, _P, _P, _P, _P],
}
_RESTYPES = {'srf_status_string': C.c_char_p, 'srf_mlp_packed_bytes': C.c_size_t,
             'srf_mlp_workspace_bytes': C.c_size_t, 'srf_compact_workspace_bytes': C.c_size_t,
             'srf_color_packed_bytes': C.c_size_t, 'srf_color_workspace_bytes': C.c_size_t,
             'srf_tune_get': C.c_int64, 'srf_tv_workspace_bytes': C.c_size_t,
             'srf_vm_bwd_workspace_bytes': C.c_size_t}

_lib = None


class SrfError(RuntimeError):
    pass


def lib():
    """Load (once) and return the ctypes handle of libsrf_hip.so."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise SrfError(f'{LIB_PATH} is missing: build it with `python -m srf_hip.build` '
                       '(the HIP kernels are the only implementation; there is no fallback)')
    handle = C.CDLL(str(LIB_PATH))
    for name, args in _PROTOS.items():
        fn = getattr(handle, name)
        fn.argtypes = args
        fn.restype = _RESTYPES.get(name, C.c_int)
    ver = handle.srf_abi_version()
    if ver != ABI_VERSION:
        raise SrfError(f'libsrf_hip.so ABI version {ver} != expected {ABI_VERSION}; rebuild it')
    _lib = handle
    return _lib


def exported_symbols() -> list[str]:
    return list(_PROTOS)


def call(name: str, *args) -> None:
    """Invoke an entry point and raise SrfError on a non-zero status."""
    fn = getattr(lib(), name)
    status = fn(*args)
    if status != 0:
        msg = lib().srf_status_string(status).decode()
        raise SrfError(f'{name} failed with status {status}: {msg}')


def ptr(t: torch.Tensor | None):
    """Device pointer of a tensor (None -> NULL)."""
    return None if t is None else t.data_ptr()


def stream_ptr(device: torch.device | None = None):
    return torch.cuda.current_stream(device).cuda_stream
