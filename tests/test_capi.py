"""CPU: the C-ABI library loads, exports every symbol include/pdc.h declares,
and without a gfx950 device refuses to run (there is no CPU fallback)."""
import ctypes
import os
import re

import pytest

from primal_dual_contacts_miror_amd import pdc
from primal_dual_contacts_miror_amd.params import PdcParams, default_params

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    txt = open(os.path.join(ROOT, "include", "pdc.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(pdc_[a-z_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    L = pdc.load_library()
    names = _declared()
    assert len(names) >= 17
    for n in names:
        assert hasattr(L, n), n
    assert set(names) == set(pdc.EXPORTS)


def test_default_params_are_reference_defaults():
    L = pdc.load_library()
    p = PdcParams()
    L.pdc_default_params(ctypes.byref(p))
    q = default_params()
    for name, _ in PdcParams._fields_:
        a, b = getattr(p, name), getattr(q, name)
        if name == "gravity":
            assert list(a) == list(b)
        else:
            assert a == b, name
    assert p.tolerance == 1e-6 and p.iterations == 100 and p.friction_coeff == 0.5
    assert p.pcg_max_iter == 1000 and p.timestep == 1e-2


def test_no_device_no_fallback():
    if pdc.device_count() > 0:
        pytest.skip("a gfx950 device is present")
    with pytest.raises(pdc.PdcError) as e:
        pdc.Context(default_params())
    assert e.value.code == pdc.PDC_ERR_NO_DEVICE
    L = pdc.load_library()
    h = ctypes.c_void_p()
    assert L.pdc_create(0, None, ctypes.byref(h)) == pdc.PDC_ERR_NO_DEVICE
    assert not h.value
    # null context handling
    assert L.pdc_step(None, 1, None) == pdc.PDC_ERR_INVALID_ARGUMENT
    assert L.pdc_last_error(None) == b"null context"


def test_shapes_struct_and_null_context():
    from primal_dual_contacts_miror_amd.params import PdcShapes
    assert ctypes.sizeof(PdcShapes) == 9 * 4
    L = pdc.load_library()
    sh = PdcShapes()
    assert L.pdc_get_shapes(None, ctypes.byref(sh)) == pdc.PDC_ERR_INVALID_ARGUMENT
    assert L.pdc_set_shapes(None, ctypes.byref(sh)) == pdc.PDC_ERR_INVALID_ARGUMENT
    assert L.pdc_autotune(None, 1, ctypes.byref(sh)) == pdc.PDC_ERR_INVALID_ARGUMENT


def test_missing_library_fails_loudly(tmp_path):
    saved = pdc._lib
    try:
        pdc._lib = None
        with pytest.raises(pdc.PdcError):
            pdc.load_library(str(tmp_path / "nope.so"))
    finally:
        pdc._lib = saved
