"""ctypes binding of libsrf_hip.so (the C-ABI declared in include/srf_hip.h).

The library is loaded from this package directory.  There is no fallback: if the
library is missing or was built for another ABI version, every entry point raises.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import torch

LIB_PATH = Path(__file__).resolve().parent / 'libsrf_hip.so'
ABI_VERSION = 1

_P = C.c_void_p
_I = C.c_int
_F = C.c_float

# name -> argtypes (restype is int unless listed in _RESTYPES)
_PROTOS = {
    'srf_abi_version': [],
    'srf_status_string': [_I],
    'srf_selftest_mfma': [_I, _P, _P, _P, _P],
    'srf_selftest_lds_dma': [_P, _P, _P],
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
    'srf_mlp_backward': [_P, _P, _I, _P, _P, _P, _P, _P, _P, _P, _P],
    'srf_mlp_backward_data': [_P, _P, _I, _P, _P, _P, _P, _P, _P],
    'srf_mlp_weight_grads': [_P, _I, _P, _P, _P, _P],
}
_RESTYPES = {'srf_status_string': C.c_char_p, 'srf_mlp_packed_bytes': C.c_size_t,
             'srf_mlp_workspace_bytes': C.c_size_t}

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
