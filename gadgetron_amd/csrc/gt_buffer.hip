// gt_buffer.hip -- cuBuffer / cuSenseBuffer (include/gt_buffer.h) over the
// plan's C ABI: NC2C accumulation into cyclic oversampled slots, a running sum
// over completed slots, and the backwards FFT, deapodization and crop that
// turn it into coil images.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "../../include/gt_buffer.h"
#include "gt_common.h"

using namespace gt_host;

namespace {

unsigned blocks(long long n) { return (unsigned)((n + 255) / 256); }

__global__ void scale_samples_kernel(const float2* __restrict__ in, float2* __restrict__ out,
                                     const float* __restrict__ dcw, long long n, long long dcw_len) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float w = dcw[i % dcw_len];
    out[i] = make_float2(in[i].x * w, in[i].y * w);
}

// acc += a - b (a: the completed slot, b: the slot about to be reused, which
// is then cleared)
__global__ void roll_kernel(float2* __restrict__ acc, const float2* __restrict__ done, float2* __restrict__ next,
                            long long n) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float2 a = done[i], b = next[i];
    float2 s = acc[i];
    s.x += a.x;
    s.y += a.y;
    s.x -= b.x;
    s.y -= b.y;
    acc[i] = s;
    next[i] = make_float2(0.0f, 0.0f);
}

__global__ void fill_kernel(float2* __restrict__ out, long long n, float2 v) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = v;
}

struct Crop {
    int D;
    long long N[3], Nos[3], off[3];
    long long nimg, nos;
};

// out[n, c] = in[n + off, c] (crop of the oversampled images, (N_os - N) >> 1)
__global__ void crop_kernel(const float2* __restrict__ in, float2* __restrict__ out, Crop g, int C) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= g.nimg * C) return;
    const long long c = i / g.nimg;
    long long r = i - c * g.nimg, src = 0, st = 1;
    for (int d = 0; d < g.D; d++) {
        const long long x = r % g.N[d];
        r /= g.N[d];
        src += (x + g.off[d]) * st;
        st *= g.Nos[d];
    }
    out[i] = in[c * g.nos + src];
}

// out[n] = sum_c conj(csm[n, c]) img[n, c] (sense_utilities.cu:7-60)
__global__ void conj_sum_kernel(const float2* __restrict__ img, const float2* __restrict__ csm,
                                float2* __restrict__ out, long long nimg, int C) {
    const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= nimg) return;
    float2 a = make_float2(0.0f, 0.0f);
    for (int c = 0; c < C; c++) {
        const float2 s = csm[n + nimg * c], v = img[n + nimg * c];
        a.x += s.x * v.x + s.y * v.y;
        a.y += s.x * v.y - s.y * v.x;
    }
    out[n] = a;
}

}  // namespace

struct gt_buffer {
    int D = 0;
    size_t N[3] = {0, 0, 0}, Nos[3] = {0, 0, 0};
    float W = 0.0f;
    unsigned C = 0, cycle_length = 0, sub_cycle_length = 0;
    unsigned cur_idx = 0, cur_sub_idx = 0;
    bool acc_empty = true;
    gt_nufft_plan* plan = nullptr;
    const float* dcw = nullptr;
    size_t dcw_len = 0;
    DevBuf acc, cyc, scaled, work;

    ~gt_buffer() {
        if (plan) (void)gt_nufft_destroy(plan);
    }
    size_t nos() const {
        size_t n = 1;
        for (int d = 0; d < D; d++) n *= Nos[d];
        return n;
    }
    size_t nimg() const {
        size_t n = 1;
        for (int d = 0; d < D; d++) n *= N[d];
        return n;
    }
    size_t slot() const { return nos() * C; }  // elements of one slot
    void os_dims(size_t* d, int& nd) const {
        nd = 0;
        for (int i = 0; i < D; i++) d[nd++] = Nos[i];
        d[nd++] = C;
    }
    static void check(gt_status s) {
        if (s != GT_OK) throw gt_error(s, last_error());
    }
};

extern "C" {

gt_status gt_buffer_create(gt_buffer** buffer) {
    return guarded([&] {
        GT_REQUIRE(buffer, GT_ERR_INVALID_ARGUMENT, "null output");
        *buffer = new gt_buffer();
    });
}

gt_status gt_buffer_destroy(gt_buffer* buffer) {
    return guarded([&] { delete buffer; });
}

gt_status gt_buffer_setup(gt_buffer* b, int D, const size_t* matrix_size, const size_t* matrix_size_os, float W,
                          unsigned num_coils, unsigned num_cycles, unsigned num_sub_cycles) {
    return guarded([&] {
        GT_REQUIRE(b && matrix_size && matrix_size_os, GT_ERR_INVALID_ARGUMENT, "null argument");
        GT_REQUIRE(D >= 1 && D <= 3, GT_ERR_UNSUPPORTED, "buffer dimensionality must be 1 to 3");
        GT_REQUIRE(num_coils > 0 && num_sub_cycles > 0, GT_ERR_INVALID_ARGUMENT, "coils and frames per cycle must be positive");
        bool sizes_changed = D != b->D || W != b->W, os_changed = D != b->D;
        for (int d = 0; d < D; d++) {
            sizes_changed |= matrix_size[d] != b->N[d] || matrix_size_os[d] != b->Nos[d];
            os_changed |= matrix_size_os[d] != b->Nos[d];
        }
        const bool coils_changed = num_coils != b->C;
        const bool cycles_changed = num_cycles + 1 != b->cycle_length;
        if (!b->plan || sizes_changed) {
            if (b->plan) gt_buffer::check(gt_nufft_destroy(b->plan));
            b->plan = nullptr;
            gt_nufft_desc d{};
            d.D = D;
            d.precision = GT_FLOAT;
            for (int i = 0; i < D; i++) {
                d.N[i] = matrix_size[i];
                d.N_os[i] = matrix_size_os[i];
            }
            d.os_factor = 0.0;
            d.W = W;
            d.conv_type = GT_CONV_STANDARD;
            d.device = -1;
            gt_buffer::check(gt_nufft_create(&b->plan, &d));
        }
        b->D = D;
        for (int i = 0; i < 3; i++) {
            b->N[i] = i < D ? matrix_size[i] : 0;
            b->Nos[i] = i < D ? matrix_size_os[i] : 0;
        }
        b->W = W;
        b->C = num_coils;
        b->cycle_length = num_cycles + 1;  // one working slot besides the full ones
        b->sub_cycle_length = num_sub_cycles;
        // New slot sizes, or a new number of slots, start the accumulation
        // again: the running sum is only consistent with the slots it was
        // summed from.
        const size_t bytes = b->slot() * sizeof(float2);
        if (os_changed || coils_changed || cycles_changed || !b->acc.ptr || !b->cyc.ptr) {
            b->acc.reserve(bytes);
            b->cyc.reserve(bytes * b->cycle_length);
            GT_HIP(hipMemset(b->acc.ptr, 0, bytes));
            GT_HIP(hipMemset(b->cyc.ptr, 0, bytes * b->cycle_length));
            b->cur_idx = b->cur_sub_idx = 0;
            b->acc_empty = true;
        }
        b->cur_sub_idx %= b->sub_cycle_length;
    });
}

gt_status gt_buffer_set_dcw(gt_buffer* b, const float* dcw, size_t dcw_len) {
    return guarded([&] {
        GT_REQUIRE(b, GT_ERR_INVALID_ARGUMENT, "null buffer");
        GT_REQUIRE(!dcw || dcw_len > 0, GT_ERR_INVALID_ARGUMENT, "dcw given with zero length");
        b->dcw = dcw;
        b->dcw_len = dcw ? dcw_len : 0;
    });
}

gt_status gt_buffer_clear(gt_buffer* b, gt_stream_t stream) {
    return guarded([&] {
        GT_REQUIRE(b && b->plan, GT_ERR_INVALID_ARGUMENT, "buffer not set up");
        hipStream_t s = S(stream);
        GT_HIP(hipMemsetAsync(b->acc.ptr, 0, b->slot() * sizeof(float2), s));
        GT_HIP(hipMemsetAsync(b->cyc.ptr, 0, b->slot() * sizeof(float2) * b->cycle_length, s));
        b->cur_idx = b->cur_sub_idx = 0;
        b->acc_empty = true;
    });
}

gt_status gt_buffer_add_frame_data(gt_buffer* b, const void* samples, size_t M, const float* traj,
                                   int* cycle_completed, gt_stream_t stream) {
    return guarded([&] {
        GT_REQUIRE(b && b->plan, GT_ERR_INVALID_ARGUMENT, "buffer not set up");
        GT_REQUIRE(samples && traj && M > 0, GT_ERR_INVALID_ARGUMENT,
                   "cuBuffer::add_frame_data: illegal input pointer");
        hipStream_t s = S(stream);
        const size_t slot = b->slot();
        float2* cur = b->cyc.as<float2>() + slot * b->cur_idx;
        gt_buffer::check(gt_nufft_preprocess(b->plan, traj, M, 1, GT_PREP_NC2C, stream));
        const void* in = samples;
        if (b->dcw) {
            const long long n = (long long)(M * b->C);
            b->scaled.reserve(size_t(n) * sizeof(float2));
            hipLaunchKernelGGL(scale_samples_kernel, dim3(blocks(n)), dim3(256), 0, s,
                               static_cast<const float2*>(samples), b->scaled.as<float2>(), b->dcw, n,
                               (long long)b->dcw_len);
            GT_HIP(hipGetLastError());
            in = b->scaled.ptr;
        }
        size_t sd[2] = {M, b->C}, gd[4];
        int gn;
        b->os_dims(gd, gn);
        gt_buffer::check(gt_nufft_convolve(b->plan, in, sd, 2, cur, gd, gn, GT_CONV_NC2C, 1, stream));
        int done = 0;
        if (b->cur_sub_idx == b->sub_cycle_length - 1) {
            // the slot is complete: into the running sum; the next slot leaves
            // it and is cleared before it is refilled
            done = 1;
            const unsigned next = (b->cur_idx + 1) % b->cycle_length;
            float2* nx = b->cyc.as<float2>() + slot * next;
            hipLaunchKernelGGL(roll_kernel, dim3(blocks((long long)slot)), dim3(256), 0, s, b->acc.as<float2>(), cur,
                               nx, (long long)slot);
            GT_HIP(hipGetLastError());
            b->acc_empty = false;
            b->cur_idx = next;
        }
        b->cur_sub_idx = (b->cur_sub_idx + 1) % b->sub_cycle_length;
        if (cycle_completed) *cycle_completed = done;
    });
}

gt_status gt_buffer_get_accumulated_coil_images(gt_buffer* b, void* out, gt_stream_t stream) {
    return guarded([&] {
        GT_REQUIRE(b && b->plan, GT_ERR_INVALID_ARGUMENT, "buffer not set up");
        GT_REQUIRE(out, GT_ERR_INVALID_ARGUMENT, "null output");
        hipStream_t s = S(stream);
        const long long nout = (long long)(b->nimg() * b->C);
        if (b->acc_empty) {  // not ready: an image of ones
            hipLaunchKernelGGL(fill_kernel, dim3(blocks(nout)), dim3(256), 0, s, static_cast<float2*>(out), nout,
                               make_float2(1.0f, 0.0f));
            GT_HIP(hipGetLastError());
            return;
        }
        const size_t slot = b->slot();
        b->work.reserve(slot * sizeof(float2));
        GT_HIP(hipMemcpyAsync(b->work.ptr, b->acc.ptr, slot * sizeof(float2), hipMemcpyDeviceToDevice, s));
        size_t gd[4];
        int gn;
        b->os_dims(gd, gn);
        gt_buffer::check(gt_nufft_fft(b->plan, b->work.ptr, gd, gn, GT_FFT_BACKWARDS, 1, stream));
        gt_buffer::check(gt_nufft_deapodize(b->plan, b->work.ptr, gd, gn, 0, stream));
        Crop g{};
        g.D = b->D;
        for (int d = 0; d < b->D; d++) {
            g.N[d] = (long long)b->N[d];
            g.Nos[d] = (long long)b->Nos[d];
            g.off[d] = (long long)((b->Nos[d] - b->N[d]) >> 1);
        }
        g.nimg = (long long)b->nimg();
        g.nos = (long long)b->nos();
        hipLaunchKernelGGL(crop_kernel, dim3(blocks(nout)), dim3(256), 0, s, b->work.as<float2>(),
                           static_cast<float2*>(out), g, int(b->C));
        GT_HIP(hipGetLastError());
    });
}

gt_status gt_buffer_get_normalization_factor(const gt_buffer* b, float* factor) {
    return guarded([&] {
        GT_REQUIRE(b && factor, GT_ERR_INVALID_ARGUMENT, "null argument");
        GT_REQUIRE(b->cycle_length > 1 && b->sub_cycle_length > 0, GT_ERR_INVALID_ARGUMENT, "buffer not set up");
        *factor = 1.0f / ((float(b->cycle_length) - 1.0f) * float(b->sub_cycle_length));
    });
}

gt_status gt_buffer_get_combined_coil_image(gt_buffer* b, const void* csm, void* out, gt_stream_t stream) {
    return guarded([&] {
        GT_REQUIRE(b && b->plan, GT_ERR_INVALID_ARGUMENT, "buffer not set up");
        GT_REQUIRE(csm, GT_ERR_INVALID_ARGUMENT, "cuSenseBuffer::get_combined_coil_image: csm not set");
        GT_REQUIRE(out, GT_ERR_INVALID_ARGUMENT, "null output");
        hipStream_t s = S(stream);
        const size_t nout = b->nimg() * b->C;
        DevBuf img;
        img.reserve(nout * sizeof(float2));
        const gt_status st = gt_buffer_get_accumulated_coil_images(b, img.ptr, stream);
        if (st != GT_OK) throw gt_error(st, last_error());
        hipLaunchKernelGGL(conj_sum_kernel, dim3(blocks((long long)b->nimg())), dim3(256), 0, s,
                           img.as<float2>(), static_cast<const float2*>(csm), static_cast<float2*>(out),
                           (long long)b->nimg(), int(b->C));
        GT_HIP(hipGetLastError());
        GT_HIP(hipStreamSynchronize(s));  // img is freed on return
    });
}

}  // extern "C"
