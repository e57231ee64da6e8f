"""Python mirror of Gadgetron's cuNFFT_plan<float, D> over the C ABI.

Names and argument meaning follow toolboxes/nfft/NFFT.h:14-175 and
toolboxes/nfft/gpu/cuNFFT.h:32-131: ``NFFTPlan.make_plan`` is
``NFFT<cuNDArray,float,D>::make_plan``, and ``preprocess``, ``compute``,
``mult_MH_M``, ``convolve``, ``fft``, ``deapodize`` and the getters keep their
reference semantics.  Errors raise :class:`GTError` (a RuntimeError), as the
reference throws std::runtime_error.

Arrays passed in are device arrays: :class:`DeviceArray`, or anything with
``data_ptr()`` and ``shape`` (a torch tensor on the GPU).  Shapes are numpy C
order, i.e. reversed Gadgetron dims: an image [Nx, Ny, C] is shape (C, Ny, Nx);
a trajectory [M, F] of 2-vectors is shape (F, M, 2).
"""
import ctypes
import enum

import numpy as np

from ._lib import GTError, check, gt_nufft_desc, lib, sizes

__all__ = [
    "NFFT_fft_mode", "NFFT_conv_mode", "NFFT_comp_mode", "NFFT_prep_mode", "ConvolutionType",
    "DeviceArray", "NFFTPlan", "GTError", "device_count",
]


class NFFT_fft_mode(enum.IntEnum):
    FORWARDS = 0
    BACKWARDS = 1


class NFFT_conv_mode(enum.IntEnum):
    C2NC = 0
    NC2C = 1


class NFFT_comp_mode(enum.IntEnum):
    FORWARDS_C2NC = 0
    FORWARDS_NC2C = 1
    BACKWARDS_C2NC = 2
    BACKWARDS_NC2C = 3


class NFFT_prep_mode(enum.IntEnum):
    C2NC = 0
    NC2C = 1
    ALL = 2


class ConvolutionType(enum.IntEnum):
    STANDARD = 0
    ATOMIC = 1
    SPARSE_MATRIX = 2


def device_count():
    n = ctypes.c_int(0)
    check(lib().gt_device_count(ctypes.byref(n)))
    return n.value


class DeviceArray:
    """Minimal owning device array (the role cuNDArray plays in the reference)."""

    def __init__(self, shape, dtype):
        self.shape = tuple(int(s) for s in shape)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = ctypes.c_void_p()
        check(lib().gt_malloc(ctypes.byref(p), max(self.nbytes, 1)))
        self._ptr = p.value

    @classmethod
    def from_numpy(cls, a, dtype=None):
        a = np.ascontiguousarray(a, dtype=dtype)
        d = cls(a.shape, a.dtype)
        if d.nbytes:
            check(lib().gt_memcpy_htod(d._ptr, a.ctypes.data_as(ctypes.c_void_p), d.nbytes, None))
        return d

    @classmethod
    def zeros(cls, shape, dtype):
        d = cls(shape, dtype)
        check(lib().gt_memset(d._ptr, 0, d.nbytes, None))
        check(lib().gt_stream_synchronize(None))
        return d

    def numpy(self):
        out = np.empty(self.shape, dtype=self.dtype)
        if self.nbytes:
            check(lib().gt_memcpy_dtoh(out.ctypes.data_as(ctypes.c_void_p), self._ptr, self.nbytes, None))
        return out

    def data_ptr(self):
        return self._ptr

    def free(self):
        if getattr(self, "_ptr", None):
            lib().gt_free(self._ptr)
            self._ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def _ptr(a):
    if a is None:
        return None
    # the ABI takes dense x-fastest arrays: a strided view (e.g. a transposed
    # torch tensor) would be read with the wrong strides
    if hasattr(a, "is_contiguous") and not a.is_contiguous():
        raise GTError(3, "array must be contiguous")
    return ctypes.c_void_p(int(a.data_ptr()))


def _gt_dims(a):
    return list(reversed([int(s) for s in a.shape]))


def _numel(a):
    return int(np.prod([int(s) for s in a.shape])) if len(a.shape) else 1


def _stream(stream):
    if stream is None:
        return None
    if isinstance(stream, int):
        return ctypes.c_void_p(stream)
    return ctypes.c_void_p(int(getattr(stream, "cuda_stream", stream)))


class NFFTPlan:
    """cuNFFT_plan<float, D> (toolboxes/nfft/gpu/cuNFFT.h:52-131)."""

    def __init__(self, matrix_size, matrix_size_os=None, W=5.5, os_factor=None,
                 conv=ConvolutionType.STANDARD, device=-1, precision="float"):
        """precision "double" makes an NFFT<cuNDArray, double, D> plan (STANDARD
        only): float64 trajectories and dcw, complex128 samples and images."""
        D = len(matrix_size)
        desc = gt_nufft_desc()
        desc.D = D
        if precision not in ("float", "double"):
            raise ValueError("precision must be 'float' or 'double'")
        desc.precision = 0 if precision == "float" else 1
        self.precision = precision
        for i, n in enumerate(matrix_size):
            desc.N[i] = int(n)
        if matrix_size_os is not None:
            for i, n in enumerate(matrix_size_os):
                desc.N_os[i] = int(n)
            desc.os_factor = 0.0
        else:
            desc.os_factor = float(1.5 if os_factor is None else os_factor)
        desc.W = float(W)
        desc.conv_type = int(conv)
        desc.device = int(device)
        h = ctypes.c_void_p()
        check(lib().gt_nufft_create(ctypes.byref(h), ctypes.byref(desc)))
        self._h = h.value
        self.D = D

    @classmethod
    def make_plan(cls, matrix_size, matrix_size_os, W, conv=ConvolutionType.STANDARD):
        """NFFT<cuNDArray,float,D>::make_plan (cuNFFT.cu:206-218)."""
        return cls(matrix_size, matrix_size_os, W, conv=conv)

    def __del__(self):
        try:
            if getattr(self, "_h", None):
                lib().gt_nufft_destroy(self._h)
                self._h = None
        except Exception:
            pass

    # -- setup queries (NFFT.h:135-151) -----------------------------------
    def get_matrix_size(self):
        out = sizes([0] * self.D)
        check(lib().gt_nufft_get_matrix_size(self._h, out))
        return [int(v) for v in out[: self.D]]

    def get_matrix_size_os(self):
        out = sizes([0] * self.D)
        check(lib().gt_nufft_get_matrix_size_os(self._h, out))
        return [int(v) for v in out[: self.D]]

    def get_W(self):
        w = ctypes.c_double()
        check(lib().gt_nufft_get_W(self._h, ctypes.byref(w)))
        return w.value

    def get_num_samples_frames(self):
        M, F = ctypes.c_size_t(), ctypes.c_size_t()
        check(lib().gt_nufft_get_num_samples_frames(self._h, ctypes.byref(M), ctypes.byref(F)))
        return M.value, F.value

    def set_tiling(self, tile=None, max_samples_per_item=0):
        t = (ctypes.c_int * 3)(*([0, 0, 0] if tile is None else list(tile) + [0] * (3 - len(tile))))
        check(lib().gt_nufft_set_tiling(self._h, t, int(max_samples_per_item)))

    STAGES = ("nc2c", "c2nc", "fft", "pad_deapod", "crop_deapod")

    def profile(self, enable=True):
        """Record HIP events around each stage (see gt_nufft_profile)."""
        check(lib().gt_nufft_profile(self._h, int(bool(enable))))

    def profile_stages(self, stages):
        """Record HIP events around the named stages only (empty: stop)."""
        mask = 0
        for st in stages:
            mask |= 1 << self.STAGES.index(st)
        check(lib().gt_nufft_profile_stages(self._h, mask))

    def profile_read(self):
        """{stage: (total_ms, launches)} since the last read."""
        n = len(self.STAGES)
        ms = (ctypes.c_double * n)()
        cnt = (ctypes.c_int * n)()
        check(lib().gt_nufft_profile_read(self._h, ms, cnt, n))
        return {s: (ms[i], cnt[i]) for i, s in enumerate(self.STAGES)}

    # -- operations -------------------------------------------------------
    def preprocess(self, trajectory, mode=NFFT_prep_mode.ALL, stream=None):
        """trajectory: device array of shape (F, M, D) or (M, D), float32."""
        shp = [int(s) for s in trajectory.shape]
        if shp[-1] != self.D:
            raise GTError(1, "trajectory's last axis must hold the D coordinates")
        M = shp[-2]
        F = _numel(trajectory) // (M * self.D)
        check(lib().gt_nufft_preprocess(self._h, _ptr(trajectory), M, F, int(mode), _stream(stream)))

    def compute(self, inp, out, dcw, mode, stream=None):
        di, do = _gt_dims(inp), _gt_dims(out)
        dl = 0 if dcw is None else _numel(dcw)
        check(lib().gt_nufft_compute(self._h, _ptr(inp), sizes(di), len(di), _ptr(out), sizes(do), len(do),
                                     _ptr(dcw), dl, int(mode), _stream(stream)))

    def compute_noalloc(self, inp, out, dcw, mode, stream=None):
        """compute() without any device allocation (gt_nufft_compute_noalloc):
        the grid lives in the library's fixed 64 MiB per-GPU scratch array.
        Float 2D STANDARD plans with fused FFT lengths only; calls that overlap
        on one GPU race."""
        di, do = _gt_dims(inp), _gt_dims(out)
        dl = 0 if dcw is None else _numel(dcw)
        check(lib().gt_nufft_compute_noalloc(self._h, _ptr(inp), sizes(di), len(di), _ptr(out), sizes(do),
                                             len(do), _ptr(dcw), dl, int(mode), _stream(stream)))

    def mult_MH_M(self, inp, out, dcw=None, stream=None):
        d = _gt_dims(inp)
        dl = 0 if dcw is None else _numel(dcw)
        check(lib().gt_nufft_mult_MH_M(self._h, _ptr(inp), _ptr(out), sizes(d), len(d), _ptr(dcw), dl,
                                       _stream(stream)))

    def convolve(self, inp, out, mode, accumulate=False, stream=None):
        di, do = _gt_dims(inp), _gt_dims(out)
        check(lib().gt_nufft_convolve(self._h, _ptr(inp), sizes(di), len(di), _ptr(out), sizes(do), len(do),
                                      int(mode), int(bool(accumulate)), _stream(stream)))

    def fft(self, data, mode, do_scale=True, stream=None):
        d = _gt_dims(data)
        check(lib().gt_nufft_fft(self._h, _ptr(data), sizes(d), len(d), int(mode), int(bool(do_scale)),
                                 _stream(stream)))

    def deapodize(self, image, fourier_domain=False, stream=None):
        d = _gt_dims(image)
        check(lib().gt_nufft_deapodize(self._h, _ptr(image), sizes(d), len(d), int(bool(fourier_domain)),
                                       _stream(stream)))

    def grad_traj(self, image, grad_samples, grad_traj, dcw=None, stream=None):
        """dL/dk of y = compute(image, FORWARDS_C2NC, dcw) given grad_samples =
        2 dL/d conj(y) (torch's convention); grad_traj has the trajectory's shape."""
        di, ds = _gt_dims(image), _gt_dims(grad_samples)
        dl = 0 if dcw is None else _numel(dcw)
        check(lib().gt_nufft_grad_traj(self._h, _ptr(image), sizes(di), len(di), _ptr(grad_samples), sizes(ds),
                                       len(ds), _ptr(dcw), dl, _ptr(grad_traj), _stream(stream)))

    def synchronize(self, stream=None):
        check(lib().gt_stream_synchronize(_stream(stream)))


class NFFTOperator:
    """NFFTOperator<cuNDArray, REAL, D> (toolboxes/nfft/NFFTOperator.h:10-37,
    NFFTOperator.hpp:13-98): a linear operator over one plan and an optional
    dcw.  mult_M = compute(FORWARDS_C2NC), mult_MH = compute(BACKWARDS_NC2C),
    mult_MH_M = the plan's mult_MH_M, each with the operator's dcw;
    accumulate computes into a temporary and adds it to out (gt_array_add),
    as the reference does with operator+=."""

    def __init__(self):
        self.plan = None
        self.dcw = None
        self.domain_dimensions = []
        self.codomain_dimensions = []

    def setup(self, matrix_size, matrix_size_os, W, precision="float"):
        """NFFTOperator.hpp:84-89: make_plan(N, N_os, W)."""
        self.plan = NFFTPlan(matrix_size, matrix_size_os, W, precision=precision)

    def preprocess(self, trajectory, stream=None):
        """NFFTOperator.hpp:91-95: NFFT_prep_mode.ALL."""
        self._need_plan("preprocess")
        self.plan.preprocess(trajectory, NFFT_prep_mode.ALL, stream)

    def set_dcw(self, dcw):
        self.dcw = dcw

    def get_dcw(self):
        return self.dcw

    def get_plan(self):
        return self.plan

    def set_domain_dimensions(self, dims):
        self.domain_dimensions = [int(d) for d in dims]

    def set_codomain_dimensions(self, dims):
        self.codomain_dimensions = [int(d) for d in dims]

    def mult_M(self, inp, out, accumulate=False, stream=None):
        self._apply("mult_M", inp, out, accumulate, stream,
                    lambda dst: self.plan.compute(inp, dst, self.dcw, NFFT_comp_mode.FORWARDS_C2NC, stream))

    def mult_MH(self, inp, out, accumulate=False, stream=None):
        self._apply("mult_MH", inp, out, accumulate, stream,
                    lambda dst: self.plan.compute(inp, dst, self.dcw, NFFT_comp_mode.BACKWARDS_NC2C, stream))

    def mult_MH_M(self, inp, out, accumulate=False, stream=None):
        """NFFTOperator.hpp:59-80: refuses to run before the codomain is set."""
        if inp is not None and out is not None and not self.codomain_dimensions:
            raise GTError(1, "NFFTOperator::mult_MH_M : operator codomain dimensions not set")
        self._apply("mult_MH_M", inp, out, accumulate, stream,
                    lambda dst: self.plan.mult_MH_M(inp, dst, self.dcw, stream))

    def _need_plan(self, who):
        if self.plan is None:
            raise GTError(1, f"NFFTOperator::{who} : setup has not been called")

    def _apply(self, who, inp, out, accumulate, stream, run):
        if inp is None or out is None:
            raise GTError(1, f"NFFTOperator::{who} : 0x0 input/output not accepted")
        self._need_plan(who)
        if not accumulate:
            run(out)
            return
        dbl = self.plan.precision == "double"
        tmp = DeviceArray(out.shape, np.complex128 if dbl else np.complex64)
        run(tmp)
        check(lib().gt_array_add(_ptr(out), _ptr(tmp), 2 * _numel(tmp), 1 if dbl else 0, _stream(stream)))
        check(lib().gt_stream_synchronize(_stream(stream)))  # before tmp is freed
        tmp.free()
