// gt_b1map.hip -- estimate_b1_map (include/gt_b1map.h) without the
// correlation matrices.
//
// The reference forms the C x C correlation matrix of every pixel (C^2 images:
// 512 MB at 256^2 x 32 coils), box-smooths each of its C^2 images and runs two
// power iterations per pixel (b1_map.cu:32-108).  With x the RSS-normalised
// coil vector and S(p) = sum_{q in box(p)} x(q) x(q)^H / 7^D, the iterations are
//   v1(p) = S(p) 1    = box(x . s)(p),  s(q) = sum_c conj(x_c(q))
//   v2(p) = S(p) u(p) = sum_q x(q) (x(q)^H u(p)) / 7^D,  u = v1 / ||v1||
// so neither needs S: the first is a box filter of a C-channel image, the
// second two passes over the pixel's box per coil chunk staged in LDS (the
// 49 dot products x(q)^H u(p) of one box plane live in registers).  Memory
// stays at three C-channel images.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "../../include/gt_b1map.h"
#include "gt_common.h"

using namespace gt_host;

namespace {

constexpr int kKW = 7;               // box width (b1_map.cu:19)
constexpr int kH = kKW >> 1;
constexpr int kT = 16;               // output tile edge in x and y (one z plane per block)
constexpr int kS = kT + 2 * kH;      // staged tile edge
constexpr int kCB = 8;               // coils per staged chunk
constexpr int kTaps = kKW * kKW;     // taps of one box plane

struct BGeom {
    int X, Y, Z, D, C;
    long long P;
};

__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// conj(a) b
__device__ __forceinline__ float2 cmulc(float2 a, float2 b) {
    return make_float2(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}

// x / rss and s = sum_c conj(x / rss)   (rss_normalize, b1_map.cu:145-187)
__global__ __launch_bounds__(256) void b1_rss_kernel(const float2* __restrict__ x, float2* __restrict__ xh,
                                                     float2* __restrict__ s, BGeom g) {
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= g.P) return;
    float r = 0.0f;
    for (int c = 0; c < g.C; c++) {
        const float2 v = x[c * g.P + p];
        r += v.x * v.x + v.y * v.y;
    }
    const float inv = 1.0f / sqrtf(r);
    float2 a = make_float2(0.0f, 0.0f);
    for (int c = 0; c < g.C; c++) {
        float2 v = x[c * g.P + p];
        v.x *= inv;
        v.y *= inv;
        xh[c * g.P + p] = v;
        a.x += v.x;
        a.y -= v.y;
    }
    s[p] = a;
}

// Stage n planes of kS x kS values (plane i from src + i * P), zero outside the image.
__device__ __forceinline__ void stage_planes(float2* dst, const float2* __restrict__ src, long long P, int n,
                                             const BGeom& g, long long plane_off, int x0, int y0) {
    const int tid = threadIdx.y * kT + threadIdx.x;
    for (int i = tid; i < n * kS * kS; i += kT * kT) {
        const int c = i / (kS * kS), r = i - c * (kS * kS);
        const int gy = y0 - kH + r / kS, gx = x0 - kH + r % kS;
        float2 v = make_float2(0.0f, 0.0f);
        if (gx >= 0 && gx < g.X && gy >= 0 && gy < g.Y) v = src[c * P + plane_off + (long long)gy * g.X + gx];
        dst[i] = v;
    }
}

// v1 = box(x . s) / 7^D: the first power iteration from ones.
__global__ __launch_bounds__(kT * kT) void b1_box_kernel(const float2* __restrict__ xh, const float2* __restrict__ s,
                                                         float2* __restrict__ v, BGeom g, float scale) {
    __shared__ float2 xs[kCB * kS * kS];
    __shared__ float2 ss[kS * kS];
    const int x0 = blockIdx.x * kT, y0 = blockIdx.y * kT, z = blockIdx.z;
    const int tx = threadIdx.x, ty = threadIdx.y;
    const int x = x0 + tx, y = y0 + ty;
    const bool in = x < g.X && y < g.Y;
    const long long plane = (long long)g.X * g.Y;
    const long long p = (long long)z * plane + (long long)y * g.X + x;
    const int hz = g.D == 3 ? kH : 0;
    for (int c0 = 0; c0 < g.C; c0 += kCB) {
        const int cb = min(kCB, g.C - c0);
        float2 acc[kCB];
#pragma unroll
        for (int c = 0; c < kCB; c++) acc[c] = make_float2(0.0f, 0.0f);
        for (int dz = -hz; dz <= hz; dz++) {
            const int zq = z + dz;
            if (zq < 0 || zq >= g.Z) continue;  // uniform over the block
            __syncthreads();
            stage_planes(xs, xh + c0 * g.P, g.P, cb, g, zq * plane, x0, y0);
            stage_planes(ss, s, 0, 1, g, zq * plane, x0, y0);
            __syncthreads();
            if (!in) continue;
#pragma unroll
            for (int dy = 0; dy < kKW; dy++)
#pragma unroll
                for (int dx = 0; dx < kKW; dx++) {
                    const int o = (ty + dy) * kS + tx + dx;
                    const float2 w = ss[o];
#pragma unroll
                    for (int c = 0; c < kCB; c++) {
                        if (c < cb) {
                            const float2 a = cmul(xs[c * kS * kS + o], w);
                            acc[c].x += a.x;
                            acc[c].y += a.y;
                        }
                    }
                }
        }
        if (in) {
#pragma unroll
            for (int c = 0; c < kCB; c++)
                if (c < cb) v[(c0 + c) * g.P + p] = make_float2(acc[c].x * scale, acc[c].y * scale);
        }
    }
}

// u = v / ||v|| per pixel (extract_csm_kernel's normalisation, b1_map.cu:734-746)
__global__ __launch_bounds__(256) void b1_normalize_kernel(float2* __restrict__ v, BGeom g) {
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= g.P) return;
    float n = 0.0f;
    for (int c = 0; c < g.C; c++) {
        const float2 a = v[c * g.P + p];
        n += a.x * a.x + a.y * a.y;
    }
    const float inv = 1.0f / sqrtf(n);
    for (int c = 0; c < g.C; c++) {
        const float2 a = v[c * g.P + p];
        v[c * g.P + p] = make_float2(a.x * inv, a.y * inv);
    }
}

// v2 = sum_q x(q) (x(q)^H u(p)), unscaled: the second power iteration.  Per box
// plane: the 49 dot products over all coil chunks, then the 49-term sums per coil.
__global__ __launch_bounds__(kT * kT) void b1_power_kernel(const float2* __restrict__ xh,
                                                           const float2* __restrict__ u, float2* __restrict__ v,
                                                           BGeom g) {
    __shared__ float2 xs[kCB * kS * kS];
    const int x0 = blockIdx.x * kT, y0 = blockIdx.y * kT, z = blockIdx.z;
    const int tx = threadIdx.x, ty = threadIdx.y;
    const int x = x0 + tx, y = y0 + ty;
    const bool in = x < g.X && y < g.Y;
    const long long plane = (long long)g.X * g.Y;
    const long long p = (long long)z * plane + (long long)y * g.X + x;
    const int hz = g.D == 3 ? kH : 0;
    bool first = true;
    for (int dz = -hz; dz <= hz; dz++) {
        const int zq = z + dz;
        if (zq < 0 || zq >= g.Z) continue;  // uniform over the block
        float2 d[kTaps];
#pragma unroll
        for (int q = 0; q < kTaps; q++) d[q] = make_float2(0.0f, 0.0f);
        for (int c0 = 0; c0 < g.C; c0 += kCB) {
            const int cb = min(kCB, g.C - c0);
            __syncthreads();
            stage_planes(xs, xh + c0 * g.P, g.P, cb, g, zq * plane, x0, y0);
            __syncthreads();
            if (!in) continue;
            float2 uc[kCB];
#pragma unroll
            for (int c = 0; c < kCB; c++) uc[c] = c < cb ? u[(c0 + c) * g.P + p] : make_float2(0.0f, 0.0f);
#pragma unroll
            for (int dy = 0; dy < kKW; dy++)
#pragma unroll
                for (int dx = 0; dx < kKW; dx++) {
                    const int o = (ty + dy) * kS + tx + dx;
                    float2 a = d[dy * kKW + dx];
#pragma unroll
                    for (int c = 0; c < kCB; c++) {
                        if (c < cb) {
                            const float2 t = cmulc(xs[c * kS * kS + o], uc[c]);
                            a.x += t.x;
                            a.y += t.y;
                        }
                    }
                    d[dy * kKW + dx] = a;
                }
        }
        for (int c0 = 0; c0 < g.C; c0 += kCB) {
            const int cb = min(kCB, g.C - c0);
            __syncthreads();
            stage_planes(xs, xh + c0 * g.P, g.P, cb, g, zq * plane, x0, y0);
            __syncthreads();
            if (!in) continue;
#pragma unroll
            for (int c = 0; c < kCB; c++) {
                if (c < cb) {
                    float2 a = first ? make_float2(0.0f, 0.0f) : v[(c0 + c) * g.P + p];
#pragma unroll
                    for (int dy = 0; dy < kKW; dy++)
#pragma unroll
                        for (int dx = 0; dx < kKW; dx++) {
                            const float2 t = cmul(xs[c * kS * kS + (ty + dy) * kS + tx + dx], d[dy * kKW + dx]);
                            a.x += t.x;
                            a.y += t.y;
                        }
                    v[(c0 + c) * g.P + p] = a;
                }
            }
        }
        first = false;
    }
}

// csm = w conj(exp(i arg w_0)) with w = scale v / ||scale v||
// (extract_csm_kernel's last normalisation and set_phase_reference_kernel,
// b1_map.cu:734-746, 775-795), in place.
__global__ __launch_bounds__(256) void b1_final_kernel(float2* __restrict__ v, BGeom g, float scale) {
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= g.P) return;
    float n = 0.0f;
    for (int c = 0; c < g.C; c++) {
        const float2 a = v[c * g.P + p];
        const float ax = a.x * scale, ay = a.y * scale;
        n += ax * ax + ay * ay;
    }
    const float inv = 1.0f / sqrtf(n);
    const float2 a0 = v[p];
    const float angle = atan2f(a0.y * scale * inv, a0.x * scale * inv);
    float sn, cs;
    sincosf(angle, &sn, &cs);
    for (int c = 0; c < g.C; c++) {
        const float2 a = v[c * g.P + p];
        const float wx = a.x * scale * inv, wy = a.y * scale * inv;
        v[c * g.P + p] = make_float2(wx * cs + wy * sn, wy * cs - wx * sn);
    }
}

unsigned pixel_blocks(long long P) { return (unsigned)((P + 255) / 256); }

}  // namespace

extern "C" gt_status gt_estimate_b1_map(const void* data, int D, const size_t* image_dims, unsigned num_coils,
                                        int target_coils, void* csm, gt_stream_t stream) {
    return guarded([&] {
        GT_REQUIRE(data && image_dims && csm, GT_ERR_INVALID_ARGUMENT, "null argument");
        GT_REQUIRE(D == 2 || D == 3, GT_ERR_UNSUPPORTED, "estimate_b1_map:: dimensionality mismatch (D must be 2 or 3).");
        GT_REQUIRE(num_coils > 0, GT_ERR_INVALID_ARGUMENT, "estimate_b1_map:: no coils");
        BGeom g{};
        g.D = D;
        g.X = int(image_dims[0]);
        g.Y = int(image_dims[1]);
        g.Z = D == 3 ? int(image_dims[2]) : 1;
        for (int d = 0; d < D; d++)
            GT_REQUIRE(image_dims[d] >= 1 && image_dims[d] < (size_t(1) << 16), GT_ERR_UNSUPPORTED,
                       "estimate_b1_map:: image sizes must be 1 to 65535");
        g.C = (target_coils <= 0 || unsigned(target_coils) > num_coils) ? int(num_coils) : target_coils;
        g.P = (long long)g.X * g.Y * g.Z;
        GT_REQUIRE(g.P * g.C < (1LL << 40), GT_ERR_UNSUPPORTED, "estimate_b1_map:: image too large");
        hipStream_t s = S(stream);
        DevBuf xh, sum, u;
        xh.reserve(size_t(g.P) * g.C * sizeof(float2));
        sum.reserve(size_t(g.P) * sizeof(float2));
        u.reserve(size_t(g.P) * g.C * sizeof(float2));
        float2* out = static_cast<float2*>(csm);
        const float scale = 1.0f / float(D == 3 ? kKW * kKW * kKW : kKW * kKW);
        const dim3 tiles((unsigned)((g.X + kT - 1) / kT), (unsigned)((g.Y + kT - 1) / kT), (unsigned)g.Z);
        hipLaunchKernelGGL(b1_rss_kernel, dim3(pixel_blocks(g.P)), dim3(256), 0, s, static_cast<const float2*>(data),
                           xh.as<float2>(), sum.as<float2>(), g);
        GT_HIP(hipGetLastError());
        hipLaunchKernelGGL(b1_box_kernel, tiles, dim3(kT, kT), 0, s, xh.as<float2>(), sum.as<float2>(),
                           u.as<float2>(), g, scale);
        GT_HIP(hipGetLastError());
        hipLaunchKernelGGL(b1_normalize_kernel, dim3(pixel_blocks(g.P)), dim3(256), 0, s, u.as<float2>(), g);
        GT_HIP(hipGetLastError());
        hipLaunchKernelGGL(b1_power_kernel, tiles, dim3(kT, kT), 0, s, xh.as<float2>(), u.as<float2>(), out, g);
        GT_HIP(hipGetLastError());
        hipLaunchKernelGGL(b1_final_kernel, dim3(pixel_blocks(g.P)), dim3(256), 0, s, out, g, scale);
        GT_HIP(hipGetLastError());
        GT_HIP(hipStreamSynchronize(s));  // the scratch is freed on return
    });
}
