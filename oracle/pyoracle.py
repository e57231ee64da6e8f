"""ctypes front end of the oracle library (test infrastructure only).

The oracle is a CPU restatement of Gadgetron's hoNFFT path (see
honfft_oracle.cpp for the reference files it follows).  Only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg import this module;
the product library never calls it.

Arrays follow Gadgetron's layout: x fastest (Fortran order), so an image of
dims [Nx, Ny, F, C] is a numpy array of shape (C, F, Ny, Nx) in C order.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

FORWARDS_C2NC, FORWARDS_NC2C, BACKWARDS_C2NC, BACKWARDS_NC2C = 0, 1, 2, 3
PREP_C2NC, PREP_NC2C, PREP_ALL = 0, 1, 2
CTOR_NOS, CTOR_ALPHA, CTOR_NOS_CU = 0, 1, 2


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        P, S, I, D, F = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int, ctypes.c_double, ctypes.c_float
        SP = ctypes.POINTER(ctypes.c_size_t)
        L.gto_last_error.restype = ctypes.c_char_p
        L.gto_plan_create.restype = P
        L.gto_plan_create.argtypes = [I, I, SP, SP, D, I, D]
        L.gto_plan_destroy.argtypes = [P]
        L.gto_plan_info.argtypes = [P, SP, ctypes.POINTER(D)]
        L.gto_preprocess.argtypes = [P, P, S, S, I]
        L.gto_compute.argtypes = [P, P, SP, I, P, SP, I, P, S, I]
        L.gto_mult_MH_M.argtypes = [P, P, SP, I, P, P, S]
        L.gto_convolve.argtypes = [P, P, P, S, I, I]
        L.gto_fft.argtypes = [P, P, S, I, I]
        L.gto_deapodize.argtypes = [P, P, S, I]
        L.gto_get_deapod_filter.argtypes = [P, P, I]
        L.gto_conv_row.argtypes = [P, S, ctypes.POINTER(ctypes.c_uint32), P, I, ctypes.POINTER(I)]
        L.gto_centred_fft.argtypes = [I, P, SP, I, I, I]
        L.gto_plain_fft.argtypes = [I, P, SP, I, I, I]
        L.gto_fftshift.argtypes = [I, P, SP, I, I, I]
        L.gto_bessi0.restype = D
        L.gto_bessi0.argtypes = [D]
        L.gto_bessi0f.restype = F
        L.gto_bessi0f.argtypes = [F]
        L.gto_kaiser_bessel.restype = D
        L.gto_kaiser_bessel.argtypes = [D, D, D, D]
        L.gto_kaiser_besself.restype = F
        L.gto_kaiser_besself.argtypes = [F, F, F, F]
        L.gto_ndft.argtypes = [I, SP, P, S, P, P, S, I]
        L.gto_num_threads.restype = I
        L.gto_set_num_threads.argtypes = [I]
        _lib = L
    return _lib


def _sizes(seq):
    arr = (ctypes.c_size_t * max(1, len(seq)))(*[int(v) for v in seq])
    return arr


def _check(rc):
    if rc != 0:
        raise RuntimeError(lib().gto_last_error().decode())


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def gt_dims(a):
    """Gadgetron dims (x fastest) of a C-ordered numpy array."""
    return list(reversed(a.shape))


class Plan:
    """hoNFFT_plan<REAL, D> restatement (hoNFFT.h:51-55 ctors).  With Nos given,
    cu=True selects cuNFFT_plan(N, N_os, W)'s deapodization scaling (unscaled
    filter FFT, cuNFFT.cu:170-202) instead of hoNFFT_plan(N, N_os, W)'s
    (scaled, hoNFFT.cpp:150-197): the outputs then differ by sqrt(prod N_os)."""

    def __init__(self, N, W, Nos=None, alpha=None, precision="float", cu=False):
        self.D = len(N)
        self.precision = 0 if precision == "float" else 1
        self.cdt = np.complex64 if self.precision == 0 else np.complex128
        self.rdt = np.float32 if self.precision == 0 else np.float64
        if alpha is not None:
            ctor, nos = CTOR_ALPHA, [0] * self.D
        else:
            ctor, nos, alpha = (CTOR_NOS_CU if cu else CTOR_NOS), list(Nos), 0.0
        h = lib().gto_plan_create(self.D, self.precision, _sizes(N), _sizes(nos), float(W), ctor, float(alpha))
        if not h:
            raise RuntimeError(lib().gto_last_error().decode())
        self.h = h
        self.N = list(N)
        Nos_out = _sizes([0] * self.D)
        beta = (ctypes.c_double * self.D)()
        _check(lib().gto_plan_info(self.h, Nos_out, beta))
        self.Nos = [int(v) for v in Nos_out[: self.D]]
        self.beta = [float(b) for b in beta]
        self.W = W
        self.M = self.F = 0

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.gto_plan_destroy(self.h)
            self.h = None

    def preprocess(self, traj, mode=PREP_ALL):
        """traj: array (..., M, D) in [-1/2, 1/2]; frames are the leading dims."""
        t = np.ascontiguousarray(traj, dtype=self.rdt)
        M = t.shape[-2]
        F = t.size // (M * self.D)
        _check(lib().gto_preprocess(self.h, _ptr(t), M, F, mode))
        self.M, self.F = M, F

    def compute(self, x, out_shape, mode, dcw=None):
        x = np.ascontiguousarray(x, dtype=self.cdt)
        out = np.zeros(out_shape, dtype=self.cdt)
        d = None if dcw is None else np.ascontiguousarray(dcw, dtype=self.rdt)
        ind, outd = gt_dims(x), gt_dims(out)
        _check(lib().gto_compute(self.h, _ptr(x), _sizes(ind), len(ind), _ptr(out), _sizes(outd), len(outd),
                                 None if d is None else _ptr(d), 0 if d is None else d.size, mode))
        return out

    def mult_MH_M(self, x, dcw=None):
        x = np.ascontiguousarray(x, dtype=self.cdt)
        out = np.zeros_like(x)
        d = None if dcw is None else np.ascontiguousarray(dcw, dtype=self.rdt)
        dims = gt_dims(x)
        _check(lib().gto_mult_MH_M(self.h, _ptr(x), _sizes(dims), len(dims), _ptr(out),
                                   None if d is None else _ptr(d), 0 if d is None else d.size))
        return out

    def convolve(self, x, mode, nbatches, out=None, accumulate=False):
        """mode 0 = C2NC (grid -> samples), 1 = NC2C (samples -> grid)."""
        x = np.ascontiguousarray(x, dtype=self.cdt)
        G = int(np.prod(self.Nos))
        n_out = nbatches * (self.M if mode == 0 else G)
        if out is None:
            out = np.zeros(n_out, dtype=self.cdt)
        _check(lib().gto_convolve(self.h, _ptr(x), _ptr(out), nbatches, mode, int(accumulate)))
        return out

    def fft(self, x, nbatches, forward, do_scale=True):
        x = np.array(x, dtype=self.cdt, copy=True, order="C")
        _check(lib().gto_fft(self.h, _ptr(x), nbatches, int(forward), int(do_scale)))
        return x

    def deapodize(self, x, nbatches, fourier_domain=False):
        x = np.array(x, dtype=self.cdt, copy=True, order="C")
        _check(lib().gto_deapodize(self.h, _ptr(x), nbatches, int(fourier_domain)))
        return x

    def deapod_filter(self, fourier_domain=False):
        out = np.zeros(list(reversed(self.Nos)), dtype=self.cdt)
        _check(lib().gto_get_deapod_filter(self.h, _ptr(out), int(fourier_domain)))
        return out

    def conv_row(self, sample, cap=4096):
        idx = (ctypes.c_uint32 * cap)()
        w = np.zeros(cap, dtype=self.rdt)
        n = ctypes.c_int(0)
        _check(lib().gto_conv_row(self.h, sample, idx, _ptr(w), cap, ctypes.byref(n)))
        return np.array(idx[: n.value], dtype=np.int64), w[: n.value].copy()


def centred_fft(x, rank, forward, precision="float"):
    p = 0 if precision == "float" else 1
    x = np.array(x, dtype=np.complex64 if p == 0 else np.complex128, copy=True, order="C")
    dims = gt_dims(x)
    _check(lib().gto_centred_fft(p, _ptr(x), _sizes(dims), len(dims), rank, int(forward)))
    return x


def plain_fft(x, rank, forward, precision="float"):
    p = 0 if precision == "float" else 1
    x = np.array(x, dtype=np.complex64 if p == 0 else np.complex128, copy=True, order="C")
    dims = gt_dims(x)
    _check(lib().gto_plain_fft(p, _ptr(x), _sizes(dims), len(dims), rank, int(forward)))
    return x


def fftshift(x, rank, inverse=False, precision="float"):
    p = 0 if precision == "float" else 1
    x = np.array(x, dtype=np.complex64 if p == 0 else np.complex128, copy=True, order="C")
    dims = gt_dims(x)
    _check(lib().gto_fftshift(p, _ptr(x), _sizes(dims), len(dims), rank, int(inverse)))
    return x


def bessi0(x, precision="double"):
    return lib().gto_bessi0(float(x)) if precision == "double" else lib().gto_bessi0f(float(x))


def kaiser_bessel(u, nos, one_over_W, beta, precision="double"):
    f = lib().gto_kaiser_bessel if precision == "double" else lib().gto_kaiser_besself
    return f(u, nos, one_over_W, beta)


def ndft(traj, N, x=None, y=None, adjoint=False):
    """FP64 direct NDFT.  traj (M, D) float64; forward x (nb, *reversed(N)) -> (nb, M)."""
    t = np.ascontiguousarray(traj, dtype=np.float64)
    M, D = t.shape
    if not adjoint:
        x = np.ascontiguousarray(x, dtype=np.complex128)
        nb = x.size // int(np.prod(N))
        out = np.zeros((nb, M), dtype=np.complex128)
        _check(lib().gto_ndft(D, _sizes(N), _ptr(t), M, _ptr(x), _ptr(out), nb, 0))
    else:
        y = np.ascontiguousarray(y, dtype=np.complex128)
        nb = y.size // M
        out = np.zeros([nb] + list(reversed(N)), dtype=np.complex128)
        _check(lib().gto_ndft(D, _sizes(N), _ptr(t), M, _ptr(y), _ptr(out), nb, 1))
    return out


def num_threads():
    return lib().gto_num_threads()


def set_num_threads(n):
    lib().gto_set_num_threads(int(n))
