"""C ABI checks that need no GPU: the library loads, exports every entry point
include/gt_nufft.h declares, and rejects invalid arguments on the host with
the status codes that replace the reference's exceptions
(cuGriddingConvolution.cu:67-80, GriddingConvolution.hpp:14-30, cuNFFT.cu:220-227)."""
import ctypes
import os
import re

import pytest

import gadgetron_amd as ga
from gadgetron_amd._lib import gt_nufft_desc, lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    out = []
    for h in ("gt_nufft.h", "gt_cgsense.h", "gt_radial.h", "gt_buffer.h", "gt_b1map.h"):
        text = open(os.path.join(ROOT, "include", h)).read()
        out += re.findall(r"^\s*(?:gt_status|const char\*)\s+(gt_\w+)\s*\(", text, flags=re.M)
    return out


def test_header_and_binding_agree():
    assert declared_symbols() == ga.EXPORTED_SYMBOLS


def test_library_exports_every_declared_symbol():
    L = lib()
    for name in declared_symbols():
        assert hasattr(L, name), name


def _desc(**kw):
    d = gt_nufft_desc()
    d.D = kw.get("D", 2)
    d.precision = kw.get("precision", 0)
    for i, n in enumerate(kw.get("N", [128, 128])):
        d.N[i] = n
    for i, n in enumerate(kw.get("N_os", [192, 192])):
        d.N_os[i] = n
    d.os_factor = kw.get("os_factor", 0.0)
    d.W = kw.get("W", 5.5)
    d.conv_type = kw.get("conv_type", 0)
    d.device = -1
    return d


@pytest.mark.parametrize("kw,status", [
    (dict(W=1.5), 1),                       # width < 1.8
    (dict(W=float("nan")), 1),
    (dict(D=5), 5),
    (dict(D=0), 5),
    (dict(precision=1, conv_type=1), 5),    # double plans are STANDARD only
    (dict(precision=2), 5),                 # no such precision
    (dict(conv_type=7), 1),                 # not a ConvolutionType
    (dict(N_os=[100, 192]), 1),             # N_os < N
    (dict(N_os=[193, 192]), 5),             # odd N_os
    (dict(os_factor=0.5), 1),               # oversampling factor < 1
    (dict(W=9.5), 5),                       # window wider than the instantiated kernels
])
def test_create_rejects_invalid_descriptors(kw, status):
    h = ctypes.c_void_p()
    rc = lib().gt_nufft_create(ctypes.byref(h), ctypes.byref(_desc(**kw)))
    assert rc == status
    assert not h.value
    assert lib().gt_last_error()


def test_null_arguments():
    h = ctypes.c_void_p()
    assert lib().gt_nufft_create(None, ctypes.byref(_desc())) == 1
    assert lib().gt_nufft_create(ctypes.byref(h), None) == 1
    assert lib().gt_nufft_preprocess(None, None, 1, 1, 2, None) == 1
    assert lib().gt_nufft_compute(None, None, None, 0, None, None, 0, None, 0, 0, None) == 1
    assert lib().gt_nufft_destroy(None) == 0


def test_python_errors_are_runtime_errors():
    with pytest.raises(ga.GTError) as e:
        ga.NFFTPlan([128, 128], [192, 192], W=1.0)
    assert isinstance(e.value, RuntimeError)
    assert "1.8" in str(e.value)


def test_device_count_without_gpu_is_graceful():
    n = ga.device_count()
    assert n >= 0


def test_array_add_validates_without_touching_memory():
    """gt_array_add (cuNDArray operator+=): an empty add is a no-op and an
    unknown precision is refused before anything is launched."""
    L = lib()
    assert L.gt_array_add(None, None, 0, 0, None) == 0
    assert L.gt_array_add(None, None, 0, 1, None) == 0
    assert L.gt_array_add(None, None, 4, 7, None) == 1
    assert L.gt_last_error()
    assert L.gt_array_add(None, None, 4, 0, None) == 1  # null arrays with n > 0


def test_elementwise_ops_validate_without_touching_memory():
    """gt_array_sqrt / gt_array_scale / gt_array_mul_weights (the CG-SENSE
    gadget's sqrt of the weights, scalar scale and weighting of the samples):
    empty arrays are no-ops, null arrays with elements are refused."""
    L = lib()
    assert L.gt_array_sqrt(None, 0, None) == 0
    assert L.gt_array_scale(None, 0, 2.0, None) == 0
    assert L.gt_array_mul_weights(None, None, 0, 3, None) == 0
    assert L.gt_array_mul_weights(None, None, 5, 0, None) == 0
    assert L.gt_array_sqrt(None, 4, None) == 1
    assert L.gt_last_error()
    assert L.gt_array_scale(None, 4, 2.0, None) == 1
    assert L.gt_array_mul_weights(None, None, 4, 2, None) == 1
