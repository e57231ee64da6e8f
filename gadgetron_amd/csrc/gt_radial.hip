// gt_radial.hip -- on-device radial trajectories and density compensation
// (include/gt_radial.h), the step before the NUFFT path:
//   2D golden-ratio / fixed-angle / variable-angle radial trajectories and the
//   Voronoi-cell radial weights of core/gpu/radial_utilities.cu, a 3D
//   kooshball generator, and the iterative (Pipe-Menon) density compensation
//   of mri/sdc (SDC.hpp, gpu/cuSDC.cu) with the Jinc kernel of
//   nfft/ConvolutionKernel.{h,hpp}.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <vector>

#include "../../include/gt_radial.h"
#include "gt_common.h"

using namespace gt_host;

namespace {

constexpr float kPi = 3.14159265358979323846f;

// Angular steps (radial_utilities.cu:13-17), in float.
__device__ __forceinline__ float golden_step(int mode) {
    return mode == GT_GR_SMALLEST ? kPi * (3.0f - sqrtf(5.0f)) * 0.5f : kPi / ((sqrtf(5.0f) + 1.0f) * 0.5f);
}

// k = (i - S/2) / S * (cos, sin)(angle), one sample per thread.
// radial_utilities.cu:19-38 (golden ratio: angle = (profile + offset) * step + pi).
__global__ void traj_golden_kernel(float2* __restrict__ out, unsigned S, unsigned nprof, float offset, int mode) {
    const unsigned long long n = (unsigned long long)S * nprof;
    const unsigned long long t = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const float fS = (float)S;
    const float i = (float)(t % S), profile = (float)(t / S);
    float sn, cs;
    sincosf((profile + offset) * golden_step(mode) + kPi, &sn, &cs);
    out[t] = make_float2((i - fS * 0.5f) * cs / fS, (i - fS * 0.5f) * sn / fS);
}

// radial_utilities.cu:113-130 (fixed angle: angle = (p + frame / F) / P * pi + offset + pi).
__global__ void traj_fixed_kernel(float2* __restrict__ out, unsigned S, unsigned P, unsigned F, float offset) {
    const unsigned long long n = (unsigned long long)S * P * F;
    const unsigned long long t = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const float fS = (float)S;
    const float i = (float)(t % S);
    const float lprofile = (float)((t / S) % P), frame = (float)(t / ((unsigned long long)S * P));
    float sn, cs;
    sincosf((lprofile + frame * (1.0f / (float)F)) * (1.0f / (float)P) * kPi + offset + kPi, &sn, &cs);
    out[t] = make_float2((i - fS * 0.5f) * cs / fS, (i - fS * 0.5f) * sn / fS);
}

// radial_utilities.cu:67-80: every profile of frame f takes angles[f] (the
// reference kernel indexes the angles by frame only and does not use its
// angular offset).
__global__ void traj_variable_kernel(float2* __restrict__ out, const float* __restrict__ angles, unsigned S,
                                     unsigned P, unsigned F) {
    const unsigned long long n = (unsigned long long)S * P * F;
    const unsigned long long t = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const float fS = (float)S;
    const float i = (float)(t % S);
    const unsigned frame = (unsigned)(t / ((unsigned long long)S * P));
    const float a = angles[frame];
    out[t] = make_float2((i - fS * 0.5f) * cosf(a) / fS, (i - fS * 0.5f) * sinf(a) / fS);
}

// ---- radial Voronoi-cell weights (radial_utilities.cu:201-380) --------------
struct P2 {
    float x, y;
};
__device__ __forceinline__ P2 sub(P2 a, P2 b) { return {a.x - b.x, a.y - b.y}; }
__device__ __forceinline__ float nrm2(P2 a) { return a.x * a.x + a.y * a.y; }
__device__ __forceinline__ float dot2(P2 a, P2 b) { return a.x * b.x + a.y * b.y; }
__device__ __forceinline__ P2 mid(P2 a, P2 b) { return {0.5f * (a.x + b.x), 0.5f * (a.y + b.y)}; }

__device__ __forceinline__ void profile_dir(bool golden, int mode, float profile, float offset, float inv_np, float& sn,
                                            float& cs) {
    if (golden) sincosf((profile + offset) * golden_step(mode), &sn, &cs);
    else sincosf(profile * inv_np * kPi, &sn, &cs);
}

// Weight of sample idx of a profile: the area of the polygon through the
// midpoints towards its two neighbours on the profile and the closest
// (previous, centre, next) samples on either side of it among all other
// profiles (the sample itself or its mirror image on each of them).
__global__ void radial_dcw_kernel(float* __restrict__ dcw, unsigned S, unsigned nprof, float alpha, float os_inv,
                                  float offset, int golden, int mode) {
    const unsigned long long n = (unsigned long long)S * nprof;
    const unsigned long long t = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const unsigned idx = (unsigned)(t % S), prof = (unsigned)(t / S);
    const float fS = (float)S, bias = fS * 0.5f, inv_np = 1.0f / (float)nprof;
    if (idx == S / 2) {  // the centre of k-space, shared by every profile
        const float radius = (alpha * os_inv) * 0.5f;
        dcw[t] = radius * radius * kPi / (float)nprof;
        return;
    }
    const float si = (float)idx;
    const float ctr = alpha * ((si - bias) * os_inv), ctr_inv = alpha * ((fS - si - bias) * os_inv);
    const float prv = alpha * ((si - bias - 1.0f) * os_inv), prv_inv = alpha * ((fS - (si - 1.0f) - bias) * os_inv);
    const float nxt = alpha * ((si - bias + 1.0f) * os_inv), nxt_inv = alpha * ((fS - (si + 1.0f) - bias) * os_inv);
    float sn, cs;
    profile_dir(golden != 0, mode, (float)prof, offset, inv_np, sn, cs);
    const P2 normal{-sn, cs};
    const P2 pos{ctr * cs, ctr * sn};
    const P2 pprev{prv * cs, prv * sn}, pnext{nxt * cs, nxt * sn};
    P2 a[3], b[3];  // (previous, centre, next) on the normal's side and on the other side
    for (int k = 0; k < 3; k++) a[k] = b[k] = P2{FLT_MAX, FLT_MAX};
    for (unsigned i = 0; i < nprof; i++) {
        if (i == prof) continue;
        profile_dir(golden != 0, mode, (float)i, offset, inv_np, sn, cs);
        const P2 c1{ctr * cs, ctr * sn}, c2{ctr_inv * cs, ctr_inv * sn};
        const P2 q1[3] = {{prv * cs, prv * sn}, c1, {nxt * cs, nxt * sn}};
        const P2 q2[3] = {{prv_inv * cs, prv_inv * sn}, c2, {nxt_inv * cs, nxt_inv * sn}};
        if (dot2(sub(c1, pos), normal) > 0.0f) {
            if (nrm2(sub(c1, pos)) < nrm2(sub(a[1], pos))) for (int k = 0; k < 3; k++) a[k] = q1[k];
        } else {
            if (nrm2(sub(c1, pos)) < nrm2(sub(b[1], pos))) for (int k = 0; k < 3; k++) b[k] = q1[k];
        }
        if (dot2(sub(c2, pos), normal) > 0.0f) {
            if (nrm2(sub(c2, pos)) < nrm2(sub(a[1], pos))) for (int k = 0; k < 3; k++) a[k] = q2[k];
        } else {
            if (nrm2(sub(c2, pos)) < nrm2(sub(b[1], pos))) for (int k = 0; k < 3; k++) b[k] = q2[k];
        }
    }
    // polygon around pos: previous, a.prev, a.ctr, a.next, next, b.next, b.ctr, b.prev
    const P2 poly[8] = {mid(pos, pprev), mid(pos, a[0]), mid(pos, a[1]), mid(pos, a[2]),
                        mid(pos, pnext), mid(pos, b[2]), mid(pos, b[1]), mid(pos, b[0])};
    float area = 0.0f;
    for (int k = 0; k < 8; k++) {
        const P2 u = poly[k], v = poly[(k + 1) & 7];
        area += u.x * v.y - v.x * u.y;
    }
    area *= 0.5f;
    dcw[t] = area < 0.0f ? -area : area;
}

// ---- 3D kooshball ----------------------------------------------------------
// Spoke directions on a generalised spiral over the sphere (equal-area rings
// in cos(theta), a longitude step of 3.6 / sqrt(n (1 - h^2))), optionally
// rotated by a 3 x 3 matrix; k = (i - S/2) / S * direction, clipped to
// [-1/2, 1/2].  The longitudes accumulate serially: one thread walks them.
__global__ void kooshball_dirs_kernel(float* __restrict__ phi, unsigned n) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    double p = 0.0;
    phi[0] = 0.0f;
    for (unsigned k = 1; k < n; k++) {
        const double h = -1.0 + 2.0 * (k + 0.5) / n;
        p = fmod(p + 3.6 / sqrt(n * (1.0 - h * h)), 2.0 * 3.14159265358979323846);
        phi[k] = (float)p;
    }
}

struct Rot {
    float m[9];
};

__global__ void kooshball_kernel(float* __restrict__ out, const float* __restrict__ phi, unsigned S, unsigned n,
                                 Rot R) {
    const unsigned long long t = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (unsigned long long)S * n) return;
    const unsigned k = (unsigned)(t / S), i = (unsigned)(t % S);
    const double h = -1.0 + 2.0 * (k + 0.5) / n;
    const double th = acos(h);
    const double d0 = sin(th) * cos((double)phi[k]), d1 = sin(th) * sin((double)phi[k]), d2 = cos(th);
    const double r = ((double)i - S / 2.0) / S;
    for (int c = 0; c < 3; c++) {
        const double dc = R.m[3 * c] * d0 + R.m[3 * c + 1] * d1 + R.m[3 * c + 2] * d2;
        const double v = fmin(fmax(dc * r, -0.5), 0.5);
        out[3 * t + c] = (float)v;
    }
}

unsigned blocks(unsigned long long n) { return (unsigned)((n + 255) / 256); }

// ---- iterative density compensation (Pipe-Menon) ----------------------------
// Jinc kernel (ConvolutionKernel.h:270-375, .hpp:307-400): radius
// r0 = 0.96960938 W; weight of a cell at distance |u| (all dimensions) is the
// degree-5 polynomial fit in t = |u| / r0, zero from t >= 0.96960938, clipped
// at 0; cells further than r0 along any axis are outside the window.
constexpr float kJincNorm = 0.96960938f;
constexpr int kSdcMaxD = 3;

struct SdcGeom {
    int D;
    int nos[kSdcMaxD];
    float nos_f[kSdcMaxD];
    int tile[kSdcMaxD];
    int ntiles[kSdcMaxD];
    int tile_cells;
    float r0;  // kernel radius in grid cells
};

__device__ __forceinline__ float jinc_weight(float dist, float r0) {
    const float t = dist / r0;
    if (t >= kJincNorm) return 0.0f;
    const float v = 0.999923599661866f + t * (0.014023570134078f + t * (-2.969947516849635f +
                    t * (0.891173008769047f + t * (2.326043034495162f + t * -1.270644932677047f))));
    return v < 0.0f ? 0.0f : v;
}

__device__ __forceinline__ int wrapc(int c, int n) {
    int r = c % n;
    return r < 0 ? r + n : r;
}

// p = (k + 1/2) N_os (the NFFT's own scaling)
__device__ __forceinline__ void sdc_pos(const float* __restrict__ traj, long long j, const SdcGeom& g, float* p) {
    for (int d = 0; d < g.D; d++) p[d] = (traj[j * g.D + d] + 0.5f) * g.nos_f[d];
}

// Distinct tiles a window touches along one dimension (at most 4).
__device__ __forceinline__ int sdc_tiles1d(float p, const SdcGeom& g, int d, int (&t)[4]) {
    const int lo = (int)ceilf(p - g.r0), hi = (int)floorf(p + g.r0);
    int n = 0, last = -1;
    for (int c = lo; c <= hi; c++) {
        const int tt = wrapc(c, g.nos[d]) / g.tile[d];
        if (tt == last) continue;
        bool seen = false;
        for (int q = 0; q < n; q++) seen |= t[q] == tt;
        if (!seen && n < 4) t[n++] = tt;
        last = tt;
    }
    return n;
}

__global__ void sdc_count_kernel(const float* __restrict__ traj, long long n, SdcGeom g, uint32_t* __restrict__ cnt) {
    const long long j = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    float p[kSdcMaxD];
    sdc_pos(traj, j, g, p);
    uint32_t c = 1;
    int t[4];
    for (int d = 0; d < g.D; d++) c *= (uint32_t)sdc_tiles1d(p[d], g, d, t);
    cnt[j] = c;
}

__global__ void sdc_emit_kernel(const float* __restrict__ traj, long long n, SdcGeom g,
                                const uint32_t* __restrict__ off, uint32_t* __restrict__ keys,
                                uint32_t* __restrict__ vals) {
    const long long j = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    float p[kSdcMaxD];
    sdc_pos(traj, j, g, p);
    int t[kSdcMaxD][4], m[kSdcMaxD] = {1, 1, 1};
    for (int d = 0; d < kSdcMaxD; d++) t[d][0] = 0;
    for (int d = 0; d < g.D; d++) m[d] = sdc_tiles1d(p[d], g, d, t[d]);
    uint32_t o = off[j];
    for (int a = 0; a < m[2]; a++)
        for (int b = 0; b < m[1]; b++)
            for (int c = 0; c < m[0]; c++) {
                keys[o] = (uint32_t)(t[0][c] + g.ntiles[0] * (t[1][b] + g.ntiles[1] * t[2][a]));
                vals[o] = (uint32_t)j;
                o++;
            }
}

__global__ void sdc_bounds_kernel(const uint32_t* __restrict__ keys, long long n, uint32_t* __restrict__ first,
                                  uint32_t* __restrict__ last) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t k = keys[i];
    if (i == 0 || keys[i - 1] != k) first[k] = (uint32_t)i;
    if (i == n - 1 || keys[i + 1] != k) last[k] = (uint32_t)(i + 1);
}

// Distance from a sample at p to grid cell c (wrapped coordinates): the
// nearest periodic image.  Returns a value > r0 when the cell is outside the
// window along some axis.
__device__ __forceinline__ float sdc_dist(const float* p, const int* c, const SdcGeom& g) {
    float s = 0.0f;
    for (int d = 0; d < g.D; d++) {
        float u = (float)c[d] - p[d];
        const float h = 0.5f * g.nos_f[d];
        if (u > h) u -= g.nos_f[d];
        else if (u < -h) u += g.nos_f[d];
        if (fabsf(u) > g.r0) return 2.0f * g.r0;
        s += u * u;
    }
    return sqrtf(s);
}

// NC2C, split for radial sampling: the tiles at the centre of k-space hold far
// more sample-tile pairs than the rest (every spoke crosses them), so a tile's
// pairs are cut into segments of kSdcSeg (a compact list: seg_off[t] is tile
// t's first segment), one 64-thread block per segment, one tile cell per
// thread.  A block stages kSdcStage pairs at a time in LDS (position and
// weight, one coalesced load per thread) and every thread walks them in
// ascending order; sdc_reduce_kernel adds a tile's segment sums in segment
// order, so the result is bitwise repeatable.
constexpr uint32_t kSdcSeg = 512;
constexpr int kSdcStage = 64;

__device__ __forceinline__ void sdc_tile_origin(int tile, const SdcGeom& g, int* org) {
    int tt = tile;
    for (int d = 0; d < g.D; d++) {
        org[d] = (tt % g.ntiles[d]) * g.tile[d];
        tt /= g.ntiles[d];
    }
}

// Segments per tile (at least one, so every tile writes its cells).
__global__ void sdc_nseg_kernel(const uint32_t* __restrict__ first, const uint32_t* __restrict__ last, uint32_t n,
                                uint32_t* __restrict__ nseg) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) nseg[t] = max(1u, (last[t] - first[t] + kSdcSeg - 1) / kSdcSeg);
}

__global__ void sdc_seg_map_kernel(const uint32_t* __restrict__ nseg, const uint32_t* __restrict__ seg_off,
                                   uint32_t n, uint32_t* __restrict__ seg_tile) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    for (uint32_t k = 0; k < nseg[t]; k++) seg_tile[seg_off[t] + k] = t;
}

__global__ __launch_bounds__(64) void sdc_nc2c_kernel(const float* __restrict__ traj, const float* __restrict__ w,
                                                      const uint32_t* __restrict__ list,
                                                      const uint32_t* __restrict__ first,
                                                      const uint32_t* __restrict__ last,
                                                      const uint32_t* __restrict__ seg_tile,
                                                      const uint32_t* __restrict__ seg_off, SdcGeom g,
                                                      float* __restrict__ partial) {
    __shared__ float4 sp[kSdcStage];
    const uint32_t sg = blockIdx.x;
    const int tile = (int)seg_tile[sg];
    const uint32_t k = sg - seg_off[tile];
    int org[kSdcMaxD] = {0, 0, 0};
    sdc_tile_origin(tile, g, org);
    int c[kSdcMaxD] = {0, 0, 0}, lc = threadIdx.x;
    for (int d = 0; d < g.D; d++) {
        c[d] = org[d] + lc % g.tile[d];
        lc /= g.tile[d];
    }
    const uint32_t b = first[tile] + k * kSdcSeg;
    const uint32_t e = min(last[tile], b + kSdcSeg);
    float acc = 0.0f;
    for (uint32_t base = b; base < e; base += kSdcStage) {
        const uint32_t i = base + threadIdx.x;
        __syncthreads();
        if (i < e) {
            const uint32_t j = list[i];
            float p[kSdcMaxD] = {0.0f, 0.0f, 0.0f};
            sdc_pos(traj, j, g, p);
            sp[threadIdx.x] = make_float4(p[0], p[1], p[2], w[j]);
        }
        __syncthreads();
        const int n = (int)min((uint32_t)kSdcStage, e - base);
        for (int q = 0; q < n; q++) {
            const float4 v = sp[q];
            const float pq[kSdcMaxD] = {v.x, v.y, v.z};
            acc += jinc_weight(sdc_dist(pq, c, g), g.r0) * v.w;
        }
    }
    partial[(long long)sg * 64 + threadIdx.x] = acc;
}

// grid cell = sum of its tile's segment sums, in segment order.
__global__ __launch_bounds__(64) void sdc_reduce_kernel(const float* __restrict__ partial,
                                                        const uint32_t* __restrict__ nseg,
                                                        const uint32_t* __restrict__ seg_off, SdcGeom g,
                                                        float* __restrict__ grid) {
    const int tile = blockIdx.x;
    int org[kSdcMaxD] = {0, 0, 0};
    sdc_tile_origin(tile, g, org);
    int lc = threadIdx.x;
    bool inside = threadIdx.x < g.tile_cells;
    long long cell = 0, st = 1;
    for (int d = 0; d < g.D; d++) {
        const int cd = org[d] + lc % g.tile[d];
        lc /= g.tile[d];
        inside &= cd < g.nos[d];
        cell += (long long)cd * st;
        st *= g.nos[d];
    }
    if (!inside) return;
    const float* src = partial + (long long)seg_off[tile] * 64 + threadIdx.x;
    const uint32_t n = nseg[tile];
    float acc = 0.0f;
    for (uint32_t q = 0; q < n; q++) acc += src[(long long)q * 64];
    grid[cell] = acc;
}

// C2NC fused with the weight update: dcw_j <- dcw_j / sum_cells w grid (0 where
// the sum is 0, cuSDC.cu safe_divides).
__global__ void sdc_c2nc_update_kernel(const float* __restrict__ traj, long long n, SdcGeom g,
                                       const float* __restrict__ grid, float* __restrict__ dcw) {
    const long long j = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    float p[kSdcMaxD];
    sdc_pos(traj, j, g, p);
    int lo[kSdcMaxD] = {0, 0, 0}, hi[kSdcMaxD] = {0, 0, 0};
    for (int d = 0; d < g.D; d++) {
        lo[d] = (int)ceilf(p[d] - g.r0);
        hi[d] = (int)floorf(p[d] + g.r0);
    }
    float acc = 0.0f;
    int c[kSdcMaxD];
    for (int cz = lo[2]; cz <= hi[2]; cz++)
        for (int cy = lo[1]; cy <= hi[1]; cy++)
            for (int cx = lo[0]; cx <= hi[0]; cx++) {
                const int u[kSdcMaxD] = {cx, cy, cz};
                float s = 0.0f;
                long long cell = 0, st = 1;
                for (int d = 0; d < g.D; d++) {
                    const float du = (float)u[d] - p[d];
                    s += du * du;
                    c[d] = wrapc(u[d], g.nos[d]);
                    cell += (long long)c[d] * st;
                    st *= g.nos[d];
                }
                acc += jinc_weight(sqrtf(s), g.r0) * grid[cell];
            }
    const float old = dcw[j];
    dcw[j] = acc == 0.0f ? 0.0f : old / acc;
}

__global__ void sdc_fill_kernel(float* __restrict__ w, long long n, const float* __restrict__ init) {
    const long long j = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (j < n) w[j] = init ? init[j] : 1.0f;
}

__global__ void sdc_range_kernel(const float* __restrict__ traj, long long n, int* __restrict__ flag) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && !(traj[i] >= -0.5f && traj[i] <= 0.5f)) flag[0] = 1;
}

}  // namespace

extern "C" {

gt_status gt_radial_trajectory_golden_ratio_2d(float* traj, unsigned samples_per_profile, unsigned profiles_per_frame,
                                               unsigned frames, unsigned profile_offset, int mode, gt_stream_t stream) {
    return guarded([&] {
        GT_REQUIRE(traj, GT_ERR_INVALID_ARGUMENT, "null trajectory");
        GT_REQUIRE(samples_per_profile > 0 && profiles_per_frame > 0 && frames > 0, GT_ERR_INVALID_ARGUMENT,
                   "empty trajectory");
        GT_REQUIRE(mode == GT_GR_SMALLEST || mode == GT_GR_ORIGINAL, GT_ERR_INVALID_ARGUMENT, "unknown golden-ratio step");
        const unsigned np = profiles_per_frame * frames;
        hipLaunchKernelGGL(traj_golden_kernel, dim3(blocks((unsigned long long)samples_per_profile * np)), dim3(256), 0,
                           S(stream), reinterpret_cast<float2*>(traj), samples_per_profile, np, (float)profile_offset,
                           mode);
        GT_HIP(hipGetLastError());
    });
}

gt_status gt_radial_trajectory_fixed_angle_2d(float* traj, unsigned samples_per_profile, unsigned profiles_per_frame,
                                              unsigned frames, float angular_offset, gt_stream_t stream) {
    return guarded([&] {
        GT_REQUIRE(traj, GT_ERR_INVALID_ARGUMENT, "null trajectory");
        GT_REQUIRE(samples_per_profile > 0 && profiles_per_frame > 0 && frames > 0, GT_ERR_INVALID_ARGUMENT,
                   "empty trajectory");
        const unsigned long long n = (unsigned long long)samples_per_profile * profiles_per_frame * frames;
        hipLaunchKernelGGL(traj_fixed_kernel, dim3(blocks(n)), dim3(256), 0, S(stream), reinterpret_cast<float2*>(traj),
                           samples_per_profile, profiles_per_frame, frames, angular_offset);
        GT_HIP(hipGetLastError());
    });
}

gt_status gt_radial_trajectory_variable_angle_2d(float* traj, const float* angles, unsigned samples_per_profile,
                                                 unsigned profiles_per_frame, unsigned frames, gt_stream_t stream) {
    return guarded([&] {
        GT_REQUIRE(traj && angles, GT_ERR_INVALID_ARGUMENT, "null trajectory or angles");
        GT_REQUIRE(samples_per_profile > 0 && profiles_per_frame > 0 && frames > 0, GT_ERR_INVALID_ARGUMENT,
                   "empty trajectory");
        const unsigned long long n = (unsigned long long)samples_per_profile * profiles_per_frame * frames;
        hipLaunchKernelGGL(traj_variable_kernel, dim3(blocks(n)), dim3(256), 0, S(stream),
                           reinterpret_cast<float2*>(traj), angles, samples_per_profile, profiles_per_frame, frames);
        GT_HIP(hipGetLastError());
    });
}

static void radial_dcw(float* dcw, unsigned S_, unsigned P, float alpha, float os_inv, float offset, int golden,
                       int mode, hipStream_t s) {
    GT_REQUIRE(dcw, GT_ERR_INVALID_ARGUMENT, "null dcw");
    // radial_utilities.cu:392-401
    GT_REQUIRE(P >= 4, GT_ERR_INVALID_ARGUMENT, "use at least four profiles");
    GT_REQUIRE(S_ > 0 && S_ % 2 == 0, GT_ERR_INVALID_ARGUMENT, "samples/profile must be even");
    hipLaunchKernelGGL(radial_dcw_kernel, dim3(blocks((unsigned long long)S_ * P)), dim3(256), 0, s, dcw, S_, P, alpha,
                       os_inv, offset, golden, mode);
    GT_HIP(hipGetLastError());
}

gt_status gt_radial_dcw_golden_ratio_2d(float* dcw, unsigned samples_per_profile, unsigned num_profiles, float alpha,
                                        float one_over_radial_oversampling_factor, unsigned profile_offset, int mode,
                                        gt_stream_t stream) {
    return guarded([&] {
        GT_REQUIRE(mode == GT_GR_SMALLEST || mode == GT_GR_ORIGINAL, GT_ERR_INVALID_ARGUMENT, "unknown golden-ratio step");
        radial_dcw(dcw, samples_per_profile, num_profiles, alpha, one_over_radial_oversampling_factor,
                   (float)profile_offset, 1, mode, S(stream));
    });
}

gt_status gt_radial_dcw_fixed_angle_2d(float* dcw, unsigned samples_per_profile, unsigned num_profiles, float alpha,
                                       float one_over_radial_oversampling_factor, gt_stream_t stream) {
    return guarded([&] {
        radial_dcw(dcw, samples_per_profile, num_profiles, alpha, one_over_radial_oversampling_factor, 0.0f, 0, 0,
                   S(stream));
    });
}

gt_status gt_kooshball_trajectory_3d(float* traj, unsigned samples_per_spoke, unsigned spokes, const float* rotation,
                                     gt_stream_t stream) {
    return guarded([&] {
        GT_REQUIRE(traj, GT_ERR_INVALID_ARGUMENT, "null trajectory");
        GT_REQUIRE(samples_per_spoke > 0 && spokes > 0, GT_ERR_INVALID_ARGUMENT, "empty trajectory");
        Rot R{};
        for (int i = 0; i < 9; i++) R.m[i] = rotation ? rotation[i] : (i % 4 == 0 ? 1.0f : 0.0f);
        DevBuf phi;
        phi.reserve(size_t(spokes) * sizeof(float));
        hipStream_t s = S(stream);
        hipLaunchKernelGGL(kooshball_dirs_kernel, dim3(1), dim3(64), 0, s, phi.as<float>(), spokes);
        hipLaunchKernelGGL(kooshball_kernel, dim3(blocks((unsigned long long)samples_per_spoke * spokes)), dim3(256), 0,
                           s, traj, phi.as<float>(), samples_per_spoke, spokes, R);
        GT_HIP(hipGetLastError());
        GT_HIP(hipStreamSynchronize(s));  // phi is freed on return
    });
}

gt_status gt_sdc_estimate_dcw(const float* traj, size_t n, int D, const size_t* matrix_size, float os_factor,
                              unsigned iterations, float kernel_width, const float* initial_dcw, float* dcw,
                              gt_stream_t stream) {
    return guarded([&] {
        GT_REQUIRE(traj && dcw && matrix_size, GT_ERR_INVALID_ARGUMENT, "null argument");
        GT_REQUIRE(D >= 1 && D <= kSdcMaxD, GT_ERR_UNSUPPORTED, "SDC supports 1 to 3 dimensions");
        GT_REQUIRE(n > 0 && n < (size_t(1) << 31), GT_ERR_INVALID_ARGUMENT, "sample count out of range");
        GT_REQUIRE(os_factor >= 1.0f && kernel_width > 0.0f, GT_ERR_INVALID_ARGUMENT,
                   "os_factor must be >= 1 and the kernel width positive");
        hipStream_t s = S(stream);
        SdcGeom g{};
        g.D = D;
        g.r0 = kJincNorm * kernel_width;
        g.tile_cells = 1;
        const int tdef = D == 1 ? 64 : (D == 2 ? 8 : 4);
        for (int d = 0; d < kSdcMaxD; d++) {
            if (d < D) {
                GT_REQUIRE(matrix_size[d] > 0, GT_ERR_INVALID_ARGUMENT, "empty matrix size");
                // matrix_size * os_factor, rounded up to a multiple of 32 (cuSDC.cu:25-37, SDC.hpp:52-58)
                size_t os = size_t(float(matrix_size[d]) * os_factor);
                os = (os + 31) / 32 * 32;
                GT_REQUIRE(double(os) > 2.0 * g.r0 + 1.0, GT_ERR_INVALID_ARGUMENT, "grid smaller than the kernel");
                g.nos[d] = int(os);
                g.nos_f[d] = float(os);
                g.tile[d] = tdef;
                g.ntiles[d] = int((os + tdef - 1) / tdef);
                // the binning lists at most 4 tiles per dimension for a window
                GT_REQUIRE(int(std::floor(2.0f * g.r0)) + 1 <= 3 * tdef, GT_ERR_UNSUPPORTED,
                           "kernel width too large for the SDC tiling");
            } else {
                g.nos[d] = 1;
                g.nos_f[d] = 1.0f;
                g.tile[d] = 1;
                g.ntiles[d] = 1;
            }
            g.tile_cells *= g.tile[d];
        }
        const long long N = (long long)n;
        {
            DevBuf flag;
            flag.reserve(sizeof(int));
            GT_HIP(hipMemsetAsync(flag.ptr, 0, sizeof(int), s));
            hipLaunchKernelGGL(sdc_range_kernel, dim3(blocks((unsigned long long)N * D)), dim3(256), 0, s, traj,
                               N * D, flag.as<int>());
            int h = 0;
            GT_HIP(hipMemcpyAsync(&h, flag.ptr, sizeof(int), hipMemcpyDeviceToHost, s));
            GT_HIP(hipStreamSynchronize(s));
            GT_REQUIRE(!h, GT_ERR_RANGE, "trajectory outside [-1/2, 1/2]");
        }
        // window-tile pairs, sorted by tile (stable: ascending sample order per tile)
        const uint32_t ntiles = uint32_t(g.ntiles[0]) * uint32_t(g.ntiles[1]) * uint32_t(g.ntiles[2]);
        DevBuf cnt, off, keys, vals, keys_s, list, temp, first, last, grid;
        cnt.reserve(size_t(N) * 4);
        off.reserve(size_t(N) * 4);
        hipLaunchKernelGGL(sdc_count_kernel, dim3(blocks(N)), dim3(256), 0, s, traj, N, g, cnt.as<uint32_t>());
        size_t tb = 0;
        GT_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, cnt.as<uint32_t>(), off.as<uint32_t>(), (int)N, s));
        temp.reserve(std::max<size_t>(tb, 1));
        GT_HIP(hipcub::DeviceScan::ExclusiveSum(temp.ptr, tb, cnt.as<uint32_t>(), off.as<uint32_t>(), (int)N, s));
        uint32_t lo = 0, lc = 0;
        GT_HIP(hipMemcpyAsync(&lo, off.as<uint32_t>() + N - 1, 4, hipMemcpyDeviceToHost, s));
        GT_HIP(hipMemcpyAsync(&lc, cnt.as<uint32_t>() + N - 1, 4, hipMemcpyDeviceToHost, s));
        GT_HIP(hipStreamSynchronize(s));
        const size_t np = size_t(lo) + lc;
        GT_REQUIRE(np < (size_t(1) << 31), GT_ERR_UNSUPPORTED, "too many sample-tile pairs");
        keys.reserve(np * 4);
        vals.reserve(np * 4);
        keys_s.reserve(np * 4);
        list.reserve(np * 4);
        hipLaunchKernelGGL(sdc_emit_kernel, dim3(blocks(N)), dim3(256), 0, s, traj, N, g, off.as<uint32_t>(),
                           keys.as<uint32_t>(), vals.as<uint32_t>());
        int end_bit = 1;
        while ((uint64_t(1) << end_bit) < uint64_t(ntiles)) end_bit++;
        tb = 0;
        GT_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, keys.as<uint32_t>(), keys_s.as<uint32_t>(),
                                                  vals.as<uint32_t>(), list.as<uint32_t>(), (int)np, 0, end_bit, s));
        temp.reserve(std::max<size_t>(tb, 1));
        GT_HIP(hipcub::DeviceRadixSort::SortPairs(temp.ptr, tb, keys.as<uint32_t>(), keys_s.as<uint32_t>(),
                                                  vals.as<uint32_t>(), list.as<uint32_t>(), (int)np, 0, end_bit, s));
        first.reserve(size_t(ntiles) * 4);
        last.reserve(size_t(ntiles) * 4);
        GT_HIP(hipMemsetAsync(first.ptr, 0, size_t(ntiles) * 4, s));
        GT_HIP(hipMemsetAsync(last.ptr, 0, size_t(ntiles) * 4, s));
        hipLaunchKernelGGL(sdc_bounds_kernel, dim3(blocks(np)), dim3(256), 0, s, keys_s.as<uint32_t>(), (long long)np,
                           first.as<uint32_t>(), last.as<uint32_t>());
        const size_t G = size_t(g.nos[0]) * g.nos[1] * g.nos[2];
        grid.reserve(G * sizeof(float));
        // the compact segment list
        DevBuf nseg, seg_off, seg_tile, partial;
        nseg.reserve(size_t(ntiles) * 4);
        seg_off.reserve(size_t(ntiles) * 4);
        hipLaunchKernelGGL(sdc_nseg_kernel, dim3(blocks(ntiles)), dim3(256), 0, s, first.as<uint32_t>(),
                           last.as<uint32_t>(), ntiles, nseg.as<uint32_t>());
        tb = 0;
        GT_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, nseg.as<uint32_t>(), seg_off.as<uint32_t>(), (int)ntiles, s));
        temp.reserve(std::max<size_t>(tb, 1));
        GT_HIP(hipcub::DeviceScan::ExclusiveSum(temp.ptr, tb, nseg.as<uint32_t>(), seg_off.as<uint32_t>(), (int)ntiles, s));
        uint32_t so = 0, sn = 0;
        GT_HIP(hipMemcpyAsync(&so, seg_off.as<uint32_t>() + ntiles - 1, 4, hipMemcpyDeviceToHost, s));
        GT_HIP(hipMemcpyAsync(&sn, nseg.as<uint32_t>() + ntiles - 1, 4, hipMemcpyDeviceToHost, s));
        GT_HIP(hipStreamSynchronize(s));
        const uint32_t nsegs = so + sn;
        seg_tile.reserve(size_t(nsegs) * 4);
        partial.reserve(size_t(nsegs) * 64 * sizeof(float));
        hipLaunchKernelGGL(sdc_seg_map_kernel, dim3(blocks(ntiles)), dim3(256), 0, s, nseg.as<uint32_t>(),
                           seg_off.as<uint32_t>(), ntiles, seg_tile.as<uint32_t>());
        hipLaunchKernelGGL(sdc_fill_kernel, dim3(blocks(N)), dim3(256), 0, s, dcw, N, initial_dcw);
        for (unsigned it = 0; it < iterations; it++) {
            hipLaunchKernelGGL(sdc_nc2c_kernel, dim3(nsegs), dim3(64), 0, s, traj, dcw, list.as<uint32_t>(),
                               first.as<uint32_t>(), last.as<uint32_t>(), seg_tile.as<uint32_t>(),
                               seg_off.as<uint32_t>(), g, partial.as<float>());
            hipLaunchKernelGGL(sdc_reduce_kernel, dim3(ntiles), dim3(64), 0, s, partial.as<float>(),
                               nseg.as<uint32_t>(), seg_off.as<uint32_t>(), g, grid.as<float>());
            hipLaunchKernelGGL(sdc_c2nc_update_kernel, dim3(blocks(N)), dim3(256), 0, s, traj, N, g, grid.as<float>(),
                               dcw);
        }
        GT_HIP(hipGetLastError());
        GT_HIP(hipStreamSynchronize(s));  // the work buffers are freed on return
    });
}

}  // extern "C"
