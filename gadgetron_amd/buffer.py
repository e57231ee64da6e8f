"""Frame accumulation into coil images (include/gt_buffer.h).

Mirrors cuBuffer<float, D> (toolboxes/mri/pmri/gpu/cuBuffer.h:13-57) and
cuSenseBuffer<float, D> (cuSenseBuffer.h): frames of coil samples are gridded
into cyclic oversampled k-space slots; completed slots form a running sum over
the last `num_cycles` cycles, which readout turns into coil images (backwards
FFT, deapodization, crop) or a coil-combined image.  Arrays are torch tensors
or DeviceArrays on the GPU, numpy order: samples (C, M) complex64 (coil
slowest), trajectory (M, D) float32, images (C, *N) complex64.
"""
import ctypes

import numpy as np

from ._lib import check, lib, sizes
from .nfft import DeviceArray, _numel, _ptr, _stream

__all__ = ["Buffer", "SenseBuffer"]


class Buffer:
    """cuBuffer<float, D>."""

    def __init__(self):
        h = ctypes.c_void_p()
        check(lib().gt_buffer_create(ctypes.byref(h)))
        self._h = h
        self._dcw = None
        self.matrix_size = ()
        self.num_coils = 0

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            lib().gt_buffer_destroy(h)
            self._h = None

    def setup(self, matrix_size, matrix_size_os, W, num_coils, num_cycles, num_sub_cycles):
        """cuBuffer::setup (cuBuffer.cpp:32-77)."""
        D = len(matrix_size)
        if len(matrix_size_os) != D:
            raise ValueError("matrix_size and matrix_size_os differ in dimensionality")
        check(lib().gt_buffer_setup(self._h, D, sizes(matrix_size), sizes(matrix_size_os), float(W), int(num_coils),
                                    int(num_cycles), int(num_sub_cycles)))
        self.matrix_size = tuple(int(v) for v in matrix_size)
        self.num_coils = int(num_coils)

    def set_dcw(self, dcw):
        """cuBuffer::set_dcw: kept by reference until replaced (None clears it)."""
        self._dcw = dcw
        check(lib().gt_buffer_set_dcw(self._h, _ptr(dcw) if dcw is not None else None,
                                      _numel(dcw) if dcw is not None else 0))

    def clear(self, stream=None):
        """cuBuffer::clear (cuBuffer.cpp:22-30)."""
        check(lib().gt_buffer_clear(self._h, _stream(stream)))

    def add_frame_data(self, samples, trajectory, stream=None):
        """cuBuffer::add_frame_data (cuBuffer.cpp:79-135); returns True when the
        frame completed a cycle slot."""
        if not self.num_coils:
            raise RuntimeError("cuBuffer::add_frame_data: buffer not set up")
        n = _numel(samples)
        if n % self.num_coils:
            raise ValueError("cuBuffer::add_frame_data: unexpected number of coils according to setup")
        M = n // self.num_coils
        if _numel(trajectory) != M * len(self.matrix_size):
            raise ValueError("cuBuffer::add_frame_data: trajectory and samples disagree")
        done = ctypes.c_int(0)
        check(lib().gt_buffer_add_frame_data(self._h, _ptr(samples), M, _ptr(trajectory), ctypes.byref(done),
                                             _stream(stream)))
        return bool(done.value)

    def get_accumulated_coil_images(self, out=None, stream=None):
        """cuBuffer::get_accumulated_coil_images (cuBuffer.cpp:137-171): (C, *N)."""
        shape = (self.num_coils,) + tuple(reversed(self.matrix_size))
        out = DeviceArray.zeros(shape, np.complex64) if out is None else out
        check(lib().gt_buffer_get_accumulated_coil_images(self._h, _ptr(out), _stream(stream)))
        return out

    def get_normalization_factor(self):
        """cuBuffer::get_normalization_factor: 1 / (num_cycles * num_sub_cycles)."""
        f = ctypes.c_float()
        check(lib().gt_buffer_get_normalization_factor(self._h, ctypes.byref(f)))
        return f.value


class SenseBuffer(Buffer):
    """cuSenseBuffer<float, D>: a Buffer with coil maps for coil combination."""

    def __init__(self):
        super().__init__()
        self.csm = None

    def set_csm(self, csm):
        """cuSenseBuffer::set_csm: (C, *N) complex64 coil maps."""
        self.csm = csm

    def get_combined_coil_image(self, out=None, stream=None):
        """cuSenseBuffer::get_combined_coil_image (cuSenseBuffer.cpp:18-38): (*N)."""
        if self.csm is None:
            raise RuntimeError("cuSenseBuffer::get_combined_coil_image: csm not set")
        out = DeviceArray.zeros(tuple(reversed(self.matrix_size)), np.complex64) if out is None else out
        check(lib().gt_buffer_get_combined_coil_image(self._h, _ptr(self.csm), _ptr(out), _stream(stream)))
        return out
