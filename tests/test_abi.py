"""CPU checks of the C-ABI library: it loads, exports every symbol the public header
declares, and rejects bad arguments with status codes before touching a GPU."""
import ctypes
import re
from pathlib import Path

import pytest

from srf_hip import _lib

REPO = Path(__file__).resolve().parent.parent
HEADER = REPO / 'include' / 'srf_hip.h'


def declared_symbols():
    text = HEADER.read_text()
    return sorted(set(re.findall(r'^\s*(?:int|const char\*|size_t)\s+(srf_\w+)\s*\(', text, re.M)))


def test_library_loads_and_version():
    assert _lib.lib().srf_abi_version() == _lib.ABI_VERSION


def test_exports_every_declared_symbol():
    lib = ctypes.CDLL(str(_lib.LIB_PATH))
    syms = declared_symbols()
    assert len(syms) >= 6
    for s in syms:
        assert hasattr(lib, s), s


def test_binding_covers_header():
    assert set(declared_symbols()) <= set(_lib.exported_symbols())


def test_status_strings():
    lib = _lib.lib()
    assert lib.srf_status_string(0) == b'ok'
    assert b'invalid' in lib.srf_status_string(2)


@pytest.mark.parametrize('call', [
    lambda L: L.srf_stratified_samples(-1, 64, 0., 1., 0, None, None, None),
    lambda L: L.srf_stratified_samples(4, 0, 0., 1., 0, None, None, None),
    lambda L: L.srf_sample_pdf(4, 1, 8, None, None, None, None, None, None, None, None),
    lambda L: L.srf_sample_pdf(4, 2000, 8, 1, 1, None, None, 1, None, None, None),
    lambda L: L.srf_fine_z_vals(4, 2, 8, 1, 1, None, 1, None, None),
    lambda L: L.srf_composite_fwd(4, 8, None, None, None, None, None, None, None, 1, 0, 1.0,
                                  None, None, None, None, None, None, None, None, None, None),
    lambda L: L.srf_composite_bwd(4, 5000, None, 1, 1, 1, 1, 1, 1, 1, 0, 1.0, 1, 1, 1,
                                  None, None, None, None, None, None, None, None, None, 1, 1, None),
])
def test_bad_arguments_return_status(call):
    assert call(_lib.lib()) in (2, 3)


def test_call_raises_on_error_status():
    with pytest.raises(_lib.SrfError):
        _lib.call('srf_stratified_samples', -1, 64, 0., 1., 0, None, None, None)
