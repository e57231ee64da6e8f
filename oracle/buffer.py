"""CPU restatement of the reference's frame-accumulation buffer -- test
infrastructure only (the checker of include/gt_buffer.h; the library never
calls it).

Follows toolboxes/mri/pmri/gpu/:
  cuBuffer.cpp:22-30    clear
  cuBuffer.cpp:32-77    setup (with the evident intent of its change flags)
  cuBuffer.cpp:79-135   add_frame_data: NC2C accumulation, slot roll
  cuBuffer.cpp:137-171  get_accumulated_coil_images: FFT, deapodize, crop
  cuBuffer.h:27-29      get_normalization_factor
  cuSenseBuffer.cpp:18-38 get_combined_coil_image (sum_c conj(csm) image)
over the hoNFFT restatement (oracle/pyoracle.py) for the gridding steps, with
cuNFFT's deapodization scaling (cuBuffer makes a cuNFFT plan from N and N_os,
cuBuffer.cpp:49-51).
Arrays are numpy order: samples (C, M), trajectory (M, D), images (C, *N).
"""
import numpy as np

from . import pyoracle as po


class Buffer:
    def __init__(self):
        self.plan = None
        self.dcw = None
        self.key = None

    def setup(self, N, Nos, W, C, num_cycles, num_sub_cycles):
        key = (tuple(N), tuple(Nos), float(np.float32(W)))
        if self.plan is None or key != self.key:
            self.plan = po.Plan(list(N), float(np.float32(W)), Nos=list(Nos), cu=True)
        restart = (self.key is None or tuple(Nos) != self.key[1] or C != getattr(self, "C", None)
                   or num_cycles + 1 != getattr(self, "L", None))
        self.key, self.N, self.Nos, self.C = key, list(N), list(Nos), int(C)
        self.L, self.sub = int(num_cycles) + 1, int(num_sub_cycles)
        if restart:
            self.G = int(np.prod(Nos))
            self.acc = np.zeros(self.C * self.G, np.complex64)
            self.cyc = np.zeros((self.L, self.C * self.G), np.complex64)
            self.cur = self.cur_sub = 0
            self.empty = True
        self.cur_sub %= self.sub

    def set_dcw(self, dcw):
        self.dcw = None if dcw is None else np.asarray(dcw, np.float32).reshape(-1)

    def clear(self):
        self.acc[:] = 0
        self.cyc[:] = 0
        self.cur = self.cur_sub = 0
        self.empty = True

    def add_frame_data(self, samples, traj):
        s = np.asarray(samples, np.complex64).reshape(-1)
        M = s.size // self.C
        self.plan.preprocess(np.asarray(traj, np.float32).reshape(1, M, len(self.N)), po.PREP_NC2C)
        if self.dcw is not None:
            s = s * self.dcw[np.arange(s.size) % self.dcw.size]
        self.plan.convolve(s, 1, self.C, out=self.cyc[self.cur], accumulate=True)
        done = self.cur_sub == self.sub - 1
        if done:
            self.acc += self.cyc[self.cur]
            self.empty = False
            self.cur = (self.cur + 1) % self.L
            self.acc -= self.cyc[self.cur]
            self.cyc[self.cur] = 0
        self.cur_sub = (self.cur_sub + 1) % self.sub
        return done

    def get_accumulated_coil_images(self):
        shape = [self.C] + list(reversed(self.N))
        if self.empty:
            return np.ones(shape, np.complex64)
        os_shape = [self.C] + list(reversed(self.Nos))
        x = self.plan.fft(self.acc.reshape(os_shape), self.C, forward=False, do_scale=True)
        x = self.plan.deapodize(x, self.C)
        sl = [slice(None)] + [slice((o - n) >> 1, ((o - n) >> 1) + n)
                              for n, o in zip(reversed(self.N), reversed(self.Nos))]
        return np.ascontiguousarray(x[tuple(sl)])

    def get_normalization_factor(self):
        return np.float32(1.0) / ((np.float32(self.L) - np.float32(1.0)) * np.float32(self.sub))

    def get_combined_coil_image(self, csm):
        img = self.get_accumulated_coil_images()
        return np.sum(np.conj(np.asarray(csm, np.complex64)) * img, axis=0).astype(np.complex64)
