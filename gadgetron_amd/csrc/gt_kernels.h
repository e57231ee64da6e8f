// gt_kernels.h -- HIP kernels of the non-Cartesian NFFT for gfx950.
//
// Hot path (reference kernels each one replaces):
//   c2nc_gather_kernel     NFFT_convolve_kernel (gpu/ConvolverC2NC.cuh:260-307)
//   nc2c_tile_kernel       NFFT_H_convolve_kernel + wrap_image_kernel
//                          (gpu/ConvolverNC2C_standard.cuh:298-477)
//   nc2c_reduce_kernel     (new) fixed-order sum of split dense tiles
//   nc2c_atomic_kernel     NFFT_H_atomic_convolve_kernel (gpu/ConvolverNC2C_atomic.cuh:216-251)
//   pad_deapod_kernel      pad_kernel + deapodize multiply + timeswitch + scale
//   crop_deapod_kernel     crop_kernel + deapodize multiply + timeswitch + scale
//                          (core/gpu/cuNDArray_utils.cu:213-350, gpu/cuNFFT.cu:73-92,
//                           fft/gpu/cuNDFFT.cu:6-36)
//   preprocess kernels     trajectory_scale, write_pairs_kernel, thrust sort/bounds
//                          (gpu/ConvolverNC2C_standard.cuh:25-190, cuGriddingConvolution.cu:128-188,361-422)
//
// Layouts: grid and image arrays are [cells x fastest, F, C] complex64 (float2);
// samples [M, F, C]; the scaled trajectory is [M*F][D] float.
//
// Centring.  For even N_os the centred FFT (ifftshift -> FFT -> fftshift) equals
// (-1)^(sum N_os/2) * T . FFT . T with T = multiply by (-1)^(x+y+z).  Both T
// passes are folded into neighbouring kernels: into the per-dimension gridding
// weights (the sign of a cell is a product of per-dimension signs) and into the
// pad/crop/deapodization kernels.  No separate timeswitch pass runs.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "gt_kb.h"

namespace gt {

struct GridGeom {
    int D;
    int nos[4];        // oversampled grid size per dim (1 for unused dims)
    int tile[4];       // tile edge per dim (1 for unused dims)
    int ntiles[4];     // tiles per dim
    int ntiles_total;  // product of ntiles
    int tile_cells;    // product of tile
    long long G;       // cells per grid
    float radius, one_over_W;
    float nos_f[4];
    float beta[4];
    // the same in double, for double-precision plans
    double radius_d, one_over_W_d;
    double nos_d[4];
    double beta_d[4];
};

struct WorkItem {
    uint32_t key;    // frame * ntiles_total + tile
    uint32_t begin;  // range into the tile-sorted sample list
    uint32_t end;
    int32_t partial; // -1: write the grid; kNoopItem: do nothing; else the partial-tile slot
};
// Placeholder item (padding of the per-XCD interleave, empty problems).
constexpr int32_t kNoopItem = -2;

struct RedItem {
    uint32_t key;
    int32_t first;   // first partial slot
    int32_t count;   // number of partial slots, summed in slot order
    int32_t pad_;
};

__device__ __forceinline__ int wrap_cell(int c, int n) {
    int r = c % n;
    return r < 0 ? r + n : r;
}

// Window of one sample along one dimension: cells ceil(p-r) .. floor(p+r)
// (ConvolutionMatrix.cpp:47-56, ConvolverC2NC.cuh:246-248), weights
// phi(|c - p|) with the per-dimension radius cut (ConvolutionKernel.hpp:164-170).
// Slots past the window get weight 0.  With `sign`, the weight carries the
// (-1)^cell factor of the centring trick.
template <int KW>
__device__ __forceinline__ void window1d(float p, const GridGeom& g, int d, bool sign,
                                         float (&w)[KW], int (&cell)[KW]) {
    const float r = g.radius;
    const int lo = (int)ceilf(p - r);
    const int hi = (int)floorf(p + r);
#pragma unroll
    for (int i = 0; i < KW; i++) {
        const int c = lo + i;
        const float u = fabsf((float)c - p);
        const bool ok = (c <= hi) && !(u > r);
        float wt = ok ? kb1(u, g.nos_f[d], g.one_over_W, g.beta[d]) : 0.0f;
        const int cw = wrap_cell(c, g.nos[d]);
        if (sign && (cw & 1)) wt = -wt;
        w[i] = wt;
        cell[i] = cw;
    }
}

// ---------------------------------------------------------------------------
// Preprocess: scale k in [-1/2, 1/2] to p = (k + 1/2) * N_os
// (hoGriddingConvolution.cpp:50-56) and flag out-of-range or non-finite input
// (cuGriddingConvolution.cu:137-152).
// ---------------------------------------------------------------------------
__global__ void scale_traj_kernel(const float* __restrict__ traj, float* __restrict__ out,
                                  long long n, int D, GridGeom g, int* __restrict__ flag) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int d = (int)(i % D);
    const float k = traj[i];
    if (!(k >= -0.5f && k <= 0.5f)) flag[0] = 1;
    out[i] = (k + 0.5f) * g.nos_f[d];
}

// SPARSE_MATRIX row normalisation (ConvolverNC2C_sparse.cuh:84-111): one over
// the sum of a sample's window weights.  The weight of a cell is the product of
// per-dimension weights, so the sum over the window is the product of the
// per-dimension sums.
__global__ void row_norm_kernel(const float* __restrict__ p, float* __restrict__ inv, long long n, GridGeom g) {
    const long long j = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    float prod = 1.0f;
    for (int d = 0; d < g.D; d++) {
        const float x = p[j * g.D + d];
        const int lo = (int)ceilf(x - g.radius), hi = (int)floorf(x + g.radius);
        float s = 0.0f;
        for (int c = lo; c <= hi; c++) {
            const float u = fabsf((float)c - x);
            if (!(u > g.radius)) s += kb1(u, g.nos_f[d], g.one_over_W, g.beta[d]);
        }
        prod *= s;
    }
    inv[j] = 1.0f / prod;
}

// w[c * FM + j] (or w[j] when the caller's dcw is per sample) = dcw * inv[j]:
// the caller's density weights times the SPARSE_MATRIX row normalisation.
__global__ void sparse_dcw_kernel(const float* __restrict__ dcw, long long dcw_len, const float* __restrict__ inv,
                                  float* __restrict__ w, long long FM, long long n) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const long long j = i % FM;
    w[i] = dcw[dcw_len == FM ? j : i % dcw_len] * inv[j];
}

// Distinct tiles touched by one sample's window along one dimension.
__device__ __forceinline__ int tiles1d(float p, const GridGeom& g, int d, int (&t)[4]) {
    const int lo = (int)ceilf(p - g.radius);
    const int hi = (int)floorf(p + g.radius);
    int n = 0;
    int last = -1;
    for (int c = lo; c <= hi; c++) {
        const int tt = wrap_cell(c, g.nos[d]) / g.tile[d];
        if (tt != last) {
            bool seen = false;
            for (int q = 0; q < n; q++) seen |= (t[q] == tt);
            if (!seen && n < 4) t[n++] = tt;
            last = tt;
        }
    }
    return n;
}

__global__ void count_tiles_kernel(const float* __restrict__ p, long long nsamp, GridGeom g,
                                   uint32_t* __restrict__ counts) {
    const long long j = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nsamp) return;
    int t[4][4];
    uint32_t cnt = 1;
    for (int d = 0; d < g.D; d++) cnt *= (uint32_t)tiles1d(p[j * g.D + d], g, d, t[d]);
    counts[j] = cnt;
}

// Emit one (tile key, sample) pair per touched tile, in sample order, so a
// stable sort by key keeps every tile's list in ascending sample order.
__global__ void emit_pairs_kernel(const float* __restrict__ p, long long nsamp, long long M,
                                  GridGeom g, const uint32_t* __restrict__ offsets,
                                  uint32_t* __restrict__ keys, uint32_t* __restrict__ vals) {
    const long long j = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nsamp) return;
    int t[4][4];
    int n[4] = {1, 1, 1, 1};
    for (int d = 0; d < g.D; d++) n[d] = tiles1d(p[j * g.D + d], g, d, t[d]);
    for (int d = g.D; d < 4; d++) t[d][0] = 0;
    const uint32_t frame = (uint32_t)(j / M);
    uint32_t o = offsets[j];
    for (int v = 0; v < n[3]; v++)
        for (int a = 0; a < n[2]; a++)
            for (int b = 0; b < n[1]; b++)
                for (int c = 0; c < n[0]; c++) {
                    const uint32_t tile =
                        (uint32_t)t[0][c] +
                        (uint32_t)g.ntiles[0] * ((uint32_t)t[1][b] +
                                                 (uint32_t)g.ntiles[1] * ((uint32_t)t[2][a] +
                                                                          (uint32_t)g.ntiles[2] * (uint32_t)t[3][v]));
                    keys[o] = frame * (uint32_t)g.ntiles_total + tile;
                    vals[o] = (uint32_t)j;
                    o++;
                }
}

// Entry ranges of a sorted key array: first[k] / last[k] bound the run of key k
// (both stay 0 for an absent key).  Found from the run boundaries, since
// atomic counting of sorted keys would serialise on every run.
__global__ void run_bounds_kernel(const uint32_t* __restrict__ keys, long long n, uint32_t* __restrict__ first,
                                  uint32_t* __restrict__ last) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t k = keys[i];
    if (i == 0 || keys[i - 1] != k) first[k] = (uint32_t)i;
    if (i == n - 1 || keys[i + 1] != k) last[k] = (uint32_t)(i + 1);
}

__global__ void run_counts_kernel(const uint32_t* __restrict__ first, const uint32_t* __restrict__ last,
                                  uint32_t nkeys, uint32_t* __restrict__ counts) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < nkeys) counts[k] = last[k] - first[k];
}

// ---------------------------------------------------------------------------
// C2NC: one lane per sample, all coils through register accumulators in
// chunks of CC; per-dimension weights computed once per sample and reused for
// every coil.  Output is multiplied by dcw or added to the old value.
// ---------------------------------------------------------------------------
template <int D, int KW, int CC>
__global__ __launch_bounds__(256) void c2nc_gather_kernel(
    const float2* __restrict__ grid, float2* __restrict__ out, const float* __restrict__ p,
    int M, int F, int C, GridGeom g, int sign, const float* __restrict__ dcw, long long dcw_len,
    int accumulate) {
    const long long nsamp = (long long)M * F;
    const long long j = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nsamp) return;
    const int f = (int)(j / M);

    float wx[KW], wy[KW], wz[KW], wv[KW];
    int cx[KW], cy[KW], cz[KW], cv[KW];
    window1d<KW>(p[j * D + 0], g, 0, sign != 0, wx, cx);
    if (D >= 2) window1d<KW>(p[j * D + 1], g, 1, sign != 0, wy, cy);
    if (D >= 3) window1d<KW>(p[j * D + 2], g, 2, sign != 0, wz, cz);
    if (D >= 4) window1d<KW>(p[j * D + 3], g, 3, sign != 0, wv, cv);

    const long long G = g.G;
    const long long FG = (long long)F * G;
    const float2* gb = grid + (long long)f * G;
    const long long FM = nsamp;

    for (int c0 = 0; c0 < C; c0 += CC) {
        float ar[CC], ai[CC];
#pragma unroll
        for (int q = 0; q < CC; q++) { ar[q] = 0.0f; ai[q] = 0.0f; }
        const int nv = D >= 4 ? KW : 1;
        const int nz = D >= 3 ? KW : 1;
        const int ny = D >= 2 ? KW : 1;
        for (int iv = 0; iv < nv; iv++)
        for (int iz = 0; iz < nz; iz++) {
            for (int iy = 0; iy < ny; iy++) {
#pragma unroll
                for (int ix = 0; ix < KW; ix++) {
                    float w;
                    long long cell;
                    if (D == 1) {
                        w = wx[ix];
                        cell = cx[ix];
                    } else if (D == 2) {
                        w = wx[ix] * wy[iy];
                        cell = cx[ix] + (long long)g.nos[0] * cy[iy];
                    } else if (D == 3) {
                        w = (wx[ix] * wy[iy]) * wz[iz];
                        cell = cx[ix] + (long long)g.nos[0] * (cy[iy] + (long long)g.nos[1] * cz[iz]);
                    } else {
                        w = ((wx[ix] * wy[iy]) * wz[iz]) * wv[iv];
                        cell = cx[ix] + (long long)g.nos[0] *
                                            (cy[iy] + (long long)g.nos[1] * (cz[iz] + (long long)g.nos[2] * cv[iv]));
                    }
                    if (w != 0.0f) {
#pragma unroll
                        for (int q = 0; q < CC; q++) {
                            if (c0 + q < C) {
                                const float2 v = gb[(long long)(c0 + q) * FG + cell];
                                ar[q] += v.x * w;
                                ai[q] += v.y * w;
                            }
                        }
                    }
                }
            }
        }
#pragma unroll
        for (int q = 0; q < CC; q++) {
            if (c0 + q < C) {
                const long long o = (long long)(c0 + q) * FM + j;
                float2 v = make_float2(ar[q], ai[q]);
                if (dcw) {
                    const float s = dcw[o % dcw_len];
                    v.x *= s;
                    v.y *= s;
                }
                if (accumulate) {
                    const float2 old = out[o];
                    v.x += old.x;
                    v.y += old.y;
                }
                out[o] = v;
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Trajectory gradient of y = dcw * C2NC(grid) (no counterpart in the reference:
// the derivative-of-Kaiser-Bessel kernel of the differentiable NUFFT).  For
// sample j and dimension d, with p = (k + 1/2) N_os,
//   dL/dk[j][d] = N_os[d] * sum_coils Re( conj(gy) * dcw * sum_cells dphi/dp_d * grid )
// where dphi/dp_d replaces dimension d's weight by its derivative in p and
// gy is the gradient of the real loss L with respect to the samples in the
// convention of torch (2 dL/d conj(y)).  A window's edge cells, whose
// weight jumps to zero, contribute their one-sided derivative.
// ---------------------------------------------------------------------------
template <int KW>
__device__ __forceinline__ void window1d_deriv(float p, const GridGeom& g, int d, bool sign, float (&dw)[KW]) {
    const float r = g.radius;
    const int lo = (int)ceilf(p - r);
    const int hi = (int)floorf(p + r);
#pragma unroll
    for (int i = 0; i < KW; i++) {
        const int c = lo + i;
        const float u = (float)c - p;
        const float au = fabsf(u);
        const bool ok = (c <= hi) && !(au > r);
        // d phi(|c - p|) / dp = -sign(c - p) phi'(|c - p|)
        float wt = ok ? kb1_deriv(au, g.nos_f[d], g.one_over_W, g.beta[d]) * (u > 0.0f ? -1.0f : 1.0f) : 0.0f;
        const int cw = wrap_cell(c, g.nos[d]);
        if (sign && (cw & 1)) wt = -wt;
        dw[i] = wt;
    }
}

template <int D, int KW>
__global__ __launch_bounds__(256) void traj_grad_kernel(const float2* __restrict__ grid,
                                                        const float2* __restrict__ gy,
                                                        const float* __restrict__ p, int M, int F, int C,
                                                        GridGeom g, int sign, const float* __restrict__ dcw,
                                                        long long dcw_len, float* __restrict__ out) {
    const long long nsamp = (long long)M * F;
    const long long j = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nsamp) return;
    const int f = (int)(j / M);
    float w[3][KW], dw[3][KW];
    int cell[3][KW];
#pragma unroll
    for (int d = 0; d < 3; d++) {
        if (d < D) {
            window1d<KW>(p[j * D + d], g, d, sign != 0, w[d], cell[d]);
            window1d_deriv<KW>(p[j * D + d], g, d, sign != 0, dw[d]);
        } else {
#pragma unroll
            for (int i = 0; i < KW; i++) {
                w[d][i] = i == 0 ? 1.0f : 0.0f;
                dw[d][i] = 0.0f;
                cell[d][i] = 0;
            }
        }
    }
    const long long G = g.G;
    const long long FG = (long long)F * G;
    const float2* gb = grid + (long long)f * G;
    float acc[3] = {0.0f, 0.0f, 0.0f};
    const int nz = D >= 3 ? KW : 1, ny = D >= 2 ? KW : 1;
    for (int c = 0; c < C; c++) {
        float sr[3] = {0.0f, 0.0f, 0.0f}, si[3] = {0.0f, 0.0f, 0.0f};
        for (int iz = 0; iz < nz; iz++) {
            for (int iy = 0; iy < ny; iy++) {
#pragma unroll
                for (int ix = 0; ix < KW; ix++) {
                    const long long cl = cell[0][ix] + (long long)g.nos[0] * (cell[1][iy] + (long long)g.nos[1] * cell[2][iz]);
                    const float wx = w[0][ix], wy = w[1][iy], wz = w[2][iz];
                    const float gx = dw[0][ix] * wy * wz, gyy = wx * dw[1][iy] * wz, gz = wx * wy * dw[2][iz];
                    if (gx == 0.0f && gyy == 0.0f && gz == 0.0f) continue;
                    const float2 v = gb[(long long)c * FG + cl];
                    sr[0] += gx * v.x;  si[0] += gx * v.y;
                    sr[1] += gyy * v.x; si[1] += gyy * v.y;
                    sr[2] += gz * v.x;  si[2] += gz * v.y;
                }
            }
        }
        const long long o = (long long)c * nsamp + j;
        const float2 gv = gy[o];
        const float s = dcw ? dcw[o % dcw_len] : 1.0f;
#pragma unroll
        for (int d = 0; d < 3; d++) acc[d] += s * (gv.x * sr[d] + gv.y * si[d]);  // Re(conj(gy) S)
    }
#pragma unroll
    for (int d = 0; d < 3; d++)
        if (d < D) out[j * D + d] = g.nos_f[d] * acc[d];
}

// ---------------------------------------------------------------------------
// NC2C, deterministic tile gather.  One wavefront per work item owns one tile
// of the oversampled grid (one frame, a chunk of coils) in LDS.  Its lanes
// take the tile's samples in ascending sample order, 64 at a time, and add
// weight * sample into the LDS tile with LDS float adds.  Only this wavefront
// touches the tile, so each cell's sum order is fixed by the sample order and
// the instruction order: the result is bitwise repeatable.  Samples whose
// windows cross several tiles are listed in each, and only the in-tile part of
// the window is added, so no halo exchange is needed.  Dense tiles are split
// into several work items whose partial tiles nc2c_reduce_kernel sums in slot
// order.
// ---------------------------------------------------------------------------
template <int D, int KW>
__global__ __launch_bounds__(64) void nc2c_tile_kernel(
    const float2* __restrict__ samples, float2* __restrict__ grid, float2* __restrict__ partials,
    const float* __restrict__ p, const uint32_t* __restrict__ tile_samples,
    const WorkItem* __restrict__ items, int M, int F, int C, int CC, GridGeom g, int sign,
    const float* __restrict__ dcw, long long dcw_len, int accumulate) {
    extern __shared__ float lds[];
    const WorkItem it = items[blockIdx.x];
    if (it.partial == kNoopItem) return;
    const int c0 = blockIdx.y * CC;
    const int ncc = min(CC, C - c0);
    const int TC = g.tile_cells;
    float* acc_re = lds;
    float* acc_im = lds + CC * TC;
    for (int i = threadIdx.x; i < 2 * CC * TC; i += 64) lds[i] = 0.0f;
    __syncthreads();

    const int f = (int)(it.key / (uint32_t)g.ntiles_total);
    int t = (int)(it.key % (uint32_t)g.ntiles_total);
    int org[4] = {0, 0, 0, 0};
    for (int d = 0; d < D; d++) {
        org[d] = (t % g.ntiles[d]) * g.tile[d];
        t /= g.ntiles[d];
    }
    const long long FM = (long long)F * M;

    for (uint32_t e0 = it.begin; e0 < it.end; e0 += 64) {
        const uint32_t e = e0 + threadIdx.x;
        const bool active = e < it.end;
        const uint32_t j = active ? tile_samples[e] : 0u;
        float wx[KW], wy[KW], wz[KW], wv[KW];
        int cx[KW], cy[KW], cz[KW], cv[KW];
        window1d<KW>(p[(long long)j * D + 0], g, 0, false, wx, cx);
        if (D >= 2) window1d<KW>(p[(long long)j * D + 1], g, 1, false, wy, cy);
        if (D >= 3) window1d<KW>(p[(long long)j * D + 2], g, 2, false, wz, cz);
        if (D >= 4) window1d<KW>(p[(long long)j * D + 3], g, 3, false, wv, cv);
        // local in-tile coordinates; -1 marks cells outside this tile
#pragma unroll
        for (int i = 0; i < KW; i++) {
            int l = cx[i] - org[0];
            cx[i] = (active && wx[i] != 0.0f && l >= 0 && l < g.tile[0]) ? l : -1;
            if (D >= 2) {
                l = cy[i] - org[1];
                cy[i] = (wy[i] != 0.0f && l >= 0 && l < g.tile[1]) ? l : -1;
            }
            if (D >= 3) {
                l = cz[i] - org[2];
                cz[i] = (wz[i] != 0.0f && l >= 0 && l < g.tile[2]) ? l : -1;
            }
            if (D >= 4) {
                l = cv[i] - org[3];
                cv[i] = (wv[i] != 0.0f && l >= 0 && l < g.tile[3]) ? l : -1;
            }
        }
        for (int q = 0; q < ncc; q++) {
            const long long sidx = (long long)(c0 + q) * FM + j;
            float2 s = active ? samples[sidx] : make_float2(0.0f, 0.0f);
            if (dcw && active) {
                const float dw = dcw[sidx % dcw_len];
                s.x *= dw;
                s.y *= dw;
            }
            float* are = acc_re + q * TC;
            float* aim = acc_im + q * TC;
            const int nv = D >= 4 ? KW : 1;
            const int nz = D >= 3 ? KW : 1;
            const int ny = D >= 2 ? KW : 1;
            for (int iv = 0; iv < nv; iv++)
            for (int iz = 0; iz < nz; iz++) {
                if (D >= 4 && cv[iv] < 0) break;  // the same for every iz
                if (D >= 3 && cz[iz] < 0) continue;
                for (int iy = 0; iy < ny; iy++) {
                    if (D >= 2 && cy[iy] < 0) continue;
#pragma unroll
                    for (int ix = 0; ix < KW; ix++) {
                        if (cx[ix] < 0) continue;
                        float w;
                        int lc;
                        if (D == 1) {
                            w = wx[ix];
                            lc = cx[ix];
                        } else if (D == 2) {
                            w = wx[ix] * wy[iy];
                            lc = cx[ix] + g.tile[0] * cy[iy];
                        } else if (D == 3) {
                            w = (wx[ix] * wy[iy]) * wz[iz];
                            lc = cx[ix] + g.tile[0] * (cy[iy] + g.tile[1] * cz[iz]);
                        } else {
                            w = ((wx[ix] * wy[iy]) * wz[iz]) * wv[iv];
                            lc = cx[ix] + g.tile[0] * (cy[iy] + g.tile[1] * (cz[iz] + g.tile[2] * cv[iv]));
                        }
                        atomicAdd(&are[lc], s.x * w);
                        atomicAdd(&aim[lc], s.y * w);
                    }
                }
            }
        }
    }
    __syncthreads();

    const long long G = g.G;
    for (int i = threadIdx.x; i < ncc * TC; i += 64) {
        const int q = i / TC;
        int lc = i % TC;
        float2 v = make_float2(acc_re[q * TC + lc], acc_im[q * TC + lc]);
        if (it.partial >= 0) {
            partials[((long long)it.partial * C + (c0 + q)) * TC + lc] = v;
            continue;
        }
        int gc[4] = {0, 0, 0, 0};
        bool inside = true;
        int par = 0;
        for (int d = 0; d < D; d++) {
            gc[d] = org[d] + lc % g.tile[d];
            lc /= g.tile[d];
            inside &= gc[d] < g.nos[d];
            par += gc[d];
        }
        if (!inside) continue;
        if (sign && (par & 1)) { v.x = -v.x; v.y = -v.y; }
        const long long cell =
            gc[0] + (long long)g.nos[0] * (gc[1] + (long long)g.nos[1] * (gc[2] + (long long)g.nos[2] * gc[3]));
        const long long o = ((long long)(c0 + q) * F + f) * G + cell;
        if (accumulate) {
            const float2 old = grid[o];
            v.x += old.x;
            v.y += old.y;
        }
        grid[o] = v;
    }
}

__global__ __launch_bounds__(256) void nc2c_reduce_kernel(
    const float2* __restrict__ partials, float2* __restrict__ grid, const RedItem* __restrict__ red,
    int F, int C, GridGeom g, int sign, int accumulate) {
    const RedItem it = red[blockIdx.x];
    const int TC = g.tile_cells;
    const int f = (int)(it.key / (uint32_t)g.ntiles_total);
    int t = (int)(it.key % (uint32_t)g.ntiles_total);
    int org[4] = {0, 0, 0, 0};
    for (int d = 0; d < g.D; d++) {
        org[d] = (t % g.ntiles[d]) * g.tile[d];
        t /= g.ntiles[d];
    }
    const long long G = g.G;
    for (int i = threadIdx.x; i < C * TC; i += blockDim.x) {
        const int q = i / TC;
        const int lc0 = i % TC;
        float2 v = make_float2(0.0f, 0.0f);
        for (int s = 0; s < it.count; s++) {
            const float2 a = partials[((long long)(it.first + s) * C + q) * TC + lc0];
            v.x += a.x;
            v.y += a.y;
        }
        int lc = lc0;
        int gc[4] = {0, 0, 0, 0};
        bool inside = true;
        int par = 0;
        for (int d = 0; d < g.D; d++) {
            gc[d] = org[d] + lc % g.tile[d];
            lc /= g.tile[d];
            inside &= gc[d] < g.nos[d];
            par += gc[d];
        }
        if (!inside) continue;
        if (sign && (par & 1)) { v.x = -v.x; v.y = -v.y; }
        const long long cell =
            gc[0] + (long long)g.nos[0] * (gc[1] + (long long)g.nos[1] * (gc[2] + (long long)g.nos[2] * gc[3]));
        const long long o = ((long long)q * F + f) * G + cell;
        if (accumulate) {
            const float2 old = grid[o];
            v.x += old.x;
            v.y += old.y;
        }
        grid[o] = v;
    }
}

// ---------------------------------------------------------------------------
// NC2C float-atomic scatter (ConvolverNC2C_atomic.cuh:59-94).  Kept only as a
// measured comparison: the summation order, hence the result, varies run to run.
// ---------------------------------------------------------------------------
template <int D, int KW>
__global__ __launch_bounds__(256) void nc2c_atomic_kernel(
    const float2* __restrict__ samples, float2* __restrict__ grid, const float* __restrict__ p,
    int M, int F, int C, GridGeom g, int sign, const float* __restrict__ dcw, long long dcw_len) {
    const long long nsamp = (long long)M * F;
    const long long j = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nsamp) return;
    const int f = (int)(j / M);
    float wx[KW], wy[KW], wz[KW], wv[KW];
    int cx[KW], cy[KW], cz[KW], cv[KW];
    window1d<KW>(p[j * D + 0], g, 0, sign != 0, wx, cx);
    if (D >= 2) window1d<KW>(p[j * D + 1], g, 1, sign != 0, wy, cy);
    if (D >= 3) window1d<KW>(p[j * D + 2], g, 2, sign != 0, wz, cz);
    if (D >= 4) window1d<KW>(p[j * D + 3], g, 3, sign != 0, wv, cv);
    const long long G = g.G;
    for (int q = 0; q < C; q++) {
        const long long sidx = (long long)q * nsamp + j;
        float2 s = samples[sidx];
        if (dcw) {
            const float dw = dcw[sidx % dcw_len];
            s.x *= dw;
            s.y *= dw;
        }
        float* gq = (float*)(grid + ((long long)q * F + f) * G);
        const int nv = D >= 4 ? KW : 1;
        const int nz = D >= 3 ? KW : 1;
        const int ny = D >= 2 ? KW : 1;
        for (int iv = 0; iv < nv; iv++)
        for (int iz = 0; iz < nz; iz++)
            for (int iy = 0; iy < ny; iy++)
#pragma unroll
                for (int ix = 0; ix < KW; ix++) {
                    float w;
                    long long cell;
                    if (D == 1) {
                        w = wx[ix];
                        cell = cx[ix];
                    } else if (D == 2) {
                        w = wx[ix] * wy[iy];
                        cell = cx[ix] + (long long)g.nos[0] * cy[iy];
                    } else if (D == 3) {
                        w = (wx[ix] * wy[iy]) * wz[iz];
                        cell = cx[ix] + (long long)g.nos[0] * (cy[iy] + (long long)g.nos[1] * cz[iz]);
                    } else {
                        w = ((wx[ix] * wy[iy]) * wz[iz]) * wv[iv];
                        cell = cx[ix] + (long long)g.nos[0] *
                                            (cy[iy] + (long long)g.nos[1] * (cz[iz] + (long long)g.nos[2] * cv[iv]));
                    }
                    if (w != 0.0f) {
                        unsafeAtomicAdd(gq + 2 * cell, s.x * w);
                        unsafeAtomicAdd(gq + 2 * cell + 1, s.y * w);
                    }
                }
    }
}

// ---------------------------------------------------------------------------
// Pad + deapodize + centring sign + scale, writing the whole oversampled grid
// (zeros outside the image).  `in_os` = the input is already oversampled.
// filt[d] holds the reciprocal 1-D deapodization filter of dim d (N_os[d]).
// ---------------------------------------------------------------------------
struct PadCrop {
    int N[4];
    int off[4];
    const float* filt[4];
    float scale;
    int sign;
    int in_os;
};

// Pad (or copy) + deapodize + centring sign + scale, caller layout, D <= 3.
// A 64-thread row per grid line R = ((b * Oz + z) * Oy + y): the row's band
// test, filter product and sign are worked out once, the threads run along x.
constexpr int kRowsPerBlock = 4;

__global__ __launch_bounds__(64 * kRowsPerBlock) void pad_deapod_kernel(const float2* __restrict__ img,
                                                                        float2* __restrict__ grid, long long nb,
                                                                        GridGeom g, PadCrop pc) {
    const long long R = (long long)blockIdx.x * kRowsPerBlock + threadIdx.y;
    const int Ox = g.nos[0], Oy = g.nos[1], Oz = g.nos[2], Ov = g.nos[3];
    if (R >= nb * Oy * Oz * Ov) return;
    const int y = (int)(R % Oy);
    const long long t = R / Oy;
    const int z = (int)(t % Oz);
    const long long t2 = t / Oz;
    const int u = (int)(t2 % Ov);
    const long long b = t2 / Ov;
    float2* out = grid + R * Ox;
    const int iy = pc.in_os ? y : y - pc.off[1], iz = pc.in_os ? z : z - pc.off[2];
    const int iv = pc.in_os ? u : u - pc.off[3];
    const int Nx = pc.in_os ? Ox : pc.N[0], Ny = pc.in_os ? Oy : pc.N[1], Nz = pc.in_os ? Oz : pc.N[2];
    const int Nv = pc.in_os ? Ov : pc.N[3];
    const int offx = pc.in_os ? 0 : pc.off[0];
    if (iy < 0 || iy >= Ny || iz < 0 || iz >= Nz || iv < 0 || iv >= Nv) {
        for (int x = threadIdx.x; x < Ox; x += 64) out[x] = make_float2(0.0f, 0.0f);
        return;
    }
    const float2* in = img + (((b * Nv + iv) * Nz + iz) * Ny + iy) * Nx;
    float sr = pc.scale;
    if (g.D >= 2) sr *= pc.filt[1][y];
    if (g.D >= 3) sr *= pc.filt[2][z];
    if (g.D >= 4) sr *= pc.filt[3][u];
    if (pc.sign && ((y + z + u) & 1)) sr = -sr;
    for (int x = threadIdx.x; x < Ox; x += 64) {
        const int ix = x - offx;
        float2 v = make_float2(0.0f, 0.0f);
        if (ix >= 0 && ix < Nx) {
            v = in[ix];
            const float s = sr * pc.filt[0][x] * ((pc.sign && (x & 1)) ? -1.0f : 1.0f);
            v.x *= s;
            v.y *= s;
        }
        out[x] = v;
    }
}

// Crop (or copy) + deapodize + centring sign + scale, caller layout, D <= 3:
// a 64-thread row per output line.
__global__ __launch_bounds__(64 * kRowsPerBlock) void crop_deapod_kernel(const float2* in, float2* out,
                                                                         long long nb, GridGeom g, PadCrop pc) {
    const int Ox = g.nos[0], Oy = g.nos[1], Oz = g.nos[2], Ov = g.nos[3];
    const int Nx = pc.in_os ? Ox : pc.N[0], Ny = pc.in_os ? Oy : pc.N[1], Nz = pc.in_os ? Oz : pc.N[2];
    const int Nv = pc.in_os ? Ov : pc.N[3];
    const long long R = (long long)blockIdx.x * kRowsPerBlock + threadIdx.y;
    if (R >= nb * Ny * Nz * Nv) return;
    const int iy = (int)(R % Ny);
    const long long t = R / Ny;
    const int iz = (int)(t % Nz);
    const long long t2 = t / Nz;
    const int iv = (int)(t2 % Nv);
    const long long b = t2 / Nv;
    const int y = pc.in_os ? iy : iy + pc.off[1], z = pc.in_os ? iz : iz + pc.off[2];
    const int u = pc.in_os ? iv : iv + pc.off[3];
    const int offx = pc.in_os ? 0 : pc.off[0];
    const float2* src = in + (((b * Ov + u) * Oz + z) * Oy + y) * (long long)Ox + offx;
    float2* dst = out + R * Nx;
    float sr = pc.scale;
    if (g.D >= 2) sr *= pc.filt[1][y];
    if (g.D >= 3) sr *= pc.filt[2][z];
    if (g.D >= 4) sr *= pc.filt[3][u];
    if (pc.sign && ((y + z + u) & 1)) sr = -sr;
    for (int ix = threadIdx.x; ix < Nx; ix += 64) {
        const int x = ix + offx;
        float2 v = src[ix];
        const float s = sr * pc.filt[0][x] * ((pc.sign && (x & 1)) ? -1.0f : 1.0f);
        v.x *= s;
        v.y *= s;
        dst[ix] = v;
    }
}

// y = 0 over n complex values, grid-stride: clears a grid inside code a caller
// may record into a hipGraph (no memset node there).
__global__ __launch_bounds__(256) void zero_kernel(float2* __restrict__ y, long long n) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        y[i] = make_float2(0.0f, 0.0f);
}

// y += x, grid-stride (gt_array_add).
template <class T>
__global__ __launch_bounds__(256) void add_kernel(T* __restrict__ y, const T* __restrict__ x, long long n) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        y[i] += x[i];
}

// y = sqrt(y), y *= a, x[b][i] *= w[i]: grid-stride element-wise kernels of
// gt_array_sqrt, gt_array_scale and gt_array_mul_weights.
__global__ __launch_bounds__(256) void sqrt_kernel(float* __restrict__ y, long long n) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        y[i] = sqrtf(y[i]);
}
__global__ __launch_bounds__(256) void scale_kernel(float* __restrict__ y, long long n, float a) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        y[i] *= a;
}
__global__ __launch_bounds__(256) void mul_weights_kernel(float2* __restrict__ x, const float* __restrict__ w,
                                                          long long n, long long total) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < total;
         i += (long long)gridDim.x * blockDim.x) {
        const float a = w[i % n];
        float2 v = x[i];
        x[i] = make_float2(v.x * a, v.y * a);
    }
}

__global__ __launch_bounds__(256) void sign_scale_kernel(float2* __restrict__ data, long long nb,
                                                         GridGeom g, float scale, int sign) {
    const long long G = g.G;
    const long long total = nb * G;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < total;
         i += (long long)gridDim.x * blockDim.x) {
        long long r = i % G;
        int par = 0;
        for (int d = 0; d < g.D; d++) {
            par += (int)(r % g.nos[d]);
            r /= g.nos[d];
        }
        float s = (sign && (par & 1)) ? -scale : scale;
        float2 v = data[i];
        v.x *= s;
        v.y *= s;
        data[i] = v;
    }
}

}  // namespace gt
