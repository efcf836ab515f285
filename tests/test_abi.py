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


def test_tune_knobs_validate_and_round_trip():
    """The schedule table is host state: set/get work without a GPU and reject bad values."""
    from srf_hip import autotune
    lib = _lib.lib()
    assert lib.srf_tune_get(99) == -1
    assert lib.srf_tune_set(99, 0) == 2
    for name, bad in (('dw_chunk', 1000), ('dw_chunk', 2048 + 128), ('color_dw_chunk', 131072),
                      ('gather_blocks', 10), ('gather_run', 0 - 1), ('gather_run', 1000)):
        assert lib.srf_tune_set(autotune.KNOBS[name], bad) == 2, (name, bad)
    try:
        for name, cands in autotune.CANDIDATES.items():
            for v in cands:
                autotune.set_knob(name, v)
                assert autotune.get_knob(name) == v
    finally:
        autotune.reset()
    assert all(autotune.get_knob(n) == 0 for n in autotune.KNOBS)


def test_autotuner_cache_round_trip(tmp_path):
    from srf_hip import autotune
    path = tmp_path / 'tune.json'
    t = autotune.Autotuner(path)
    t.results['w'] = {'gather_run': {'chosen': 16, 'ms': {'0': 2.0, '16': 1.0}}}
    t._save()
    try:
        assert autotune.Autotuner(path).apply('w')
        assert autotune.get_knob('gather_run') == 16
        assert not autotune.Autotuner(path).apply('other')
    finally:
        autotune.reset()


def test_color_descriptor_precisions():
    """Packed-weight sizes of the 300^2 colour network (144 features, features PE 2, first layer
    138 wide) in each operand precision: FP32 fragments pad reductions to 16 features; the BF16
    modes pad to 32 and hold one 1 KB fragment per BF16 plane (split: three forward, two
    backward).  Unsupported precisions and first layers past 240 inputs are rejected."""
    from srf_hip import color
    base = dict(n_features=144, features_pe=2)
    nb = lambda **kw: _lib.lib().srf_color_packed_bytes(ctypes.byref(color.ColorSpec(**{**base, **kw}).desc()))
    kb = 256  # floats per 1 KB fragment
    bias = 128 + 128 + 16
    # F32: basis 2 x 9, W1 8 x 9, W2 8 x 8, W3 1 x 8, W3^T 8 x 1, W2^T 8 x 8, W1^T 9 x 8, B^T 9 x 2 fragments
    f32 = (2 * 9 + 8 * 9 + 8 * 8 + 8 + 8 + 8 * 8 + 9 * 8 + 9 * 2) * kb + bias
    assert nb(prec='fp32') == 4 * f32
    # BF16 (units of 32 features): forward basis 2 x 5, W1 8 x 5, W2 8 x 4, W3 1 x 4; backward W3^T 8 x 1,
    # W2^T 8 x 4, W1^T 10 x 4, B^T 10 x 1
    fwd, bwd = 2 * 5 + 8 * 5 + 8 * 4 + 4, 8 + 8 * 4 + 10 * 4 + 10
    assert nb(prec='bf16') == 4 * ((fwd + bwd) * kb + bias)
    assert nb(prec='bf16_split') == 4 * ((3 * fwd + 2 * bwd) * kb + bias)
    # first layer 27 x 7 + 3 = 192 wide: supported; 27 x 9 + 3 = 246: past the limit
    for prec in ('fp32', 'bf16', 'bf16_split'):
        assert nb(prec=prec, features_pe=3) > 0 and nb(prec=prec, features_pe=4) == 0
    d = color.ColorSpec(**base).desc()
    d.prec = 2  # SRF_PREC_F16: no colour kernels
    assert _lib.lib().srf_color_packed_bytes(ctypes.byref(d)) == 0
    with pytest.raises(_lib.SrfError):
        color.ColorSpec(**base, prec='fp16').desc()
