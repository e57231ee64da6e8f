// gt_kernels_3d.h -- the 3D gridding kernels of the hot path.
//
// nc2c_reg3d_kernel (NFFT_H_convolve_kernel, ConvolverNC2C_standard.cuh:298-355)
//   One wavefront owns one 8 x 8 x 4 tile of the grid for 8 coils.  Lane
//   (plane z, coil, real or imaginary part) holds the tile's 8 x 8 cells of
//   plane z in 64 registers.  A table entry carries the sample's weights for
//   the 8 tile columns, 8 tile rows and 4 tile planes (zero outside the
//   window) and the mask of rows the window covers; a sample costs one
//   multiply by the lane's plane weight and 8 FMAs per covered row.  Tables
//   and samples stream through LDS two chunks ahead as in the 2D kernel, and
//   each lane sums its cells in table order (bitwise repeatable).
//
// c2nc_multi3d_kernel (NFFT_convolve_kernel, ConvolverC2NC.cuh:260-307)
//   A persistent workgroup stages one 4 x 4 x 4 tile plus the kernel halo
//   for 8 coils in LDS, one plane per coil in the grid's own order, so the
//   staging loads read whole grid rows; the next item's staging is in flight
//   while the current one is gathered.  Each wavefront gathers four samples at
//   a time: lane (sample slot, x half, coil) reads whole complex cells and
//   sums its half of every x row of the window; the halves are added across
//   lanes.  With the matrix cores off (GT_NUFFT_MFMA=0) or windows of 7 cells.
//
// nc2c_mfma3d_kernel, c2nc_mfma3d_kernel: the same two directions on the
// matrix cores (split-fp16 products, FP32 sums), described where they are
// defined below.
#pragma once

#include "gt_kernels_reg.h"

namespace gt {

// ---- NC2C ------------------------------------------------------------------
constexpr int k3T0 = 8, k3T1 = 8, k3T2 = 4;  // register tile (x, y, z)
constexpr int k3Coils = 8;                   // coils per wavefront (4 planes x 8 coils x 2)
constexpr int k3NW = 20;                     // weight floats per table entry: 8 x, 8 y, 4 z
constexpr int k3NK = 16;                     // table entries per chunk moved through LDS
constexpr int k3XsRow = 2 * k3Coils + 2;     // floats per entry of the sample transpose
constexpr int k3LoopFloats = 2 * k3NK * k3NW + 2 * 2 * k3NK + k3NK * k3XsRow;
constexpr int k3OutFloats = 32 * kSubP;      // write-out: half a plane tile (32 cells) x 64 lanes
constexpr int k3NC2CLdsFloats = ((k3OutFloats > k3LoopFloats ? k3OutFloats : k3LoopFloats) + 3) & ~3;
static_assert(k3NK * k3NW / 4 <= 2 * kLanes, "a chunk's weights load as at most two float4 per lane");
static_assert(k3T2 * 2 * k3Coils == kLanes, "lanes = planes x coils x (re, im)");

// Weight of grid cell gc (wrapped coordinate) for a sample at p along dim d:
// the window's unwrapped cell that wraps to gc, if any, weighted as in
// window_raw (the window is never longer than the grid).
__device__ __forceinline__ float cell_weight(float p, const GridGeom& g, int d, int gc) {
    const float r = g.radius;
    const int lo = (int)ceilf(p - r);
    const int hi = (int)floorf(p + r);
    const int c = lo + wrap_cell(gc - lo, g.nos[d]);
    if (c > hi) return 0.0f;
    const float u = fabsf((float)c - p);
    return u > r ? 0.0f : kb1(u, g.nos_f[d], g.one_over_W, g.beta[d]);
}

__device__ __forceinline__ void tile_origin3(int tile, const GridGeom& g, int (&org)[3]) {
    org[0] = (tile % g.ntiles[0]) * g.tile[0];
    tile /= g.ntiles[0];
    org[1] = (tile % g.ntiles[1]) * g.tile[1];
    org[2] = (tile / g.ntiles[1]) * g.tile[2];
}

// NC2C table, per (tile, sample) pair e, in two arrays:
//   wt[20 e + c]   weight of tile column c (c < 8), row c - 8 (c < 16), plane c - 16
//   hd[2 e]        sample index j
//   hd[2 e + 1]    bits 0-7: tile rows the window covers
__global__ void nc2c_table3d_kernel(const float* __restrict__ p, const uint32_t* __restrict__ list,
                                    const uint32_t* __restrict__ keys, long long n, GridGeom g,
                                    float* __restrict__ wt, uint32_t* __restrict__ hd) {
    const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const uint32_t j = list[e];
    int org[3];
    tile_origin3((int)(keys[e] % (uint32_t)g.ntiles_total), g, org);
    const float px = p[(long long)j * 3 + 0], py = p[(long long)j * 3 + 1], pz = p[(long long)j * 3 + 2];
    float w[k3NW];
    uint32_t rows = 0;
#pragma unroll
    for (int c = 0; c < k3T0; c++) w[c] = cell_weight(px, g, 0, org[0] + c);
#pragma unroll
    for (int c = 0; c < k3T1; c++) {
        w[k3T0 + c] = cell_weight(py, g, 1, org[1] + c);
        rows |= (w[k3T0 + c] != 0.0f ? 1u : 0u) << c;
    }
#pragma unroll
    for (int c = 0; c < k3T2; c++) w[k3T0 + k3T1 + c] = cell_weight(pz, g, 2, org[2] + c);
    float4* t = reinterpret_cast<float4*>(wt + e * k3NW);
#pragma unroll
    for (int q = 0; q < k3NW / 4; q++) t[q] = make_float4(w[4 * q], w[4 * q + 1], w[4 * q + 2], w[4 * q + 3]);
    hd[2 * e] = j;
    hd[2 * e + 1] = rows;
}

// Loads chunk entries [cb, cb + k3NK) of the table (entries [begin, end), end >
// begin): float4 f of the chunk's weights in lane f (and lane f - 64 for f >=
// 64), the head dwords in lanes 0..31.  Entries past `end` read as zeros
// (sample 0 with no rows).  Addresses are clamped so that every call issues
// the same loads.
__device__ __forceinline__ void nc2c3_chunk_load(const float* __restrict__ wt, const uint32_t* __restrict__ hd,
                                                 uint32_t cb, uint32_t end, int lane, float4& w0, float4& w1,
                                                 uint32_t& h) {
    constexpr int kF4 = k3NW / 4;  // float4 per entry
    const float4* w4 = reinterpret_cast<const float4*>(wt);
    const uint32_t e0 = cb + (uint32_t)(lane / kF4);
    const float4 a = w4[(size_t)min(e0, end - 1) * kF4 + lane % kF4];
    w0 = e0 < end ? a : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    const int f1 = kLanes + (lane & 15);  // lanes 16..63 repeat lanes 0..15 (only 0..15 are stored)
    const uint32_t e1 = cb + (uint32_t)(f1 / kF4);
    const float4 b = w4[(size_t)min(e1, end - 1) * kF4 + f1 % kF4];
    w1 = e1 < end ? b : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    const uint32_t eh = cb + (uint32_t)((lane & 31) >> 1);
    const uint32_t hv = hd[(size_t)min(eh, end - 1) * 2 + (lane & 1)];
    h = eh < end ? hv : 0u;
}

// DM: how dcw is applied (dcw_mode): 0 none, 1 per sample, 2 per coil element.
template <int DM>
__global__ __launch_bounds__(64, 3) void nc2c_reg3d_kernel(
    const float* __restrict__ samples, float* __restrict__ grid, float* __restrict__ partials,
    const float* __restrict__ wt, const uint32_t* __restrict__ hd, const WorkItem* __restrict__ items, int M,
    int F, int C, GridGeom g, Layout L, int sign, const float* __restrict__ dcw, long long dcw_len,
    int accumulate) {
    __shared__ float4 smem[k3NC2CLdsFloats / 4];
    float* xp = reinterpret_cast<float*>(smem);                                  // write-out transpose
    float (*sw)[k3NK * k3NW] = reinterpret_cast<float (*)[k3NK * k3NW]>(smem);  // weights of two chunks
    uint32_t (*sh)[2 * k3NK] = reinterpret_cast<uint32_t (*)[2 * k3NK]>(reinterpret_cast<float*>(smem) + 2 * k3NK * k3NW);
    float* xs = reinterpret_cast<float*>(smem) + 2 * k3NK * k3NW + 2 * 2 * k3NK;  // [entry][coil][re, im]
    const int lane = threadIdx.x;
    const WorkItem it = items[blockIdx.x];
    if (it.partial == kNoopItem) return;
    const int c0 = blockIdx.y * k3Coils;
    const long long FM = (long long)F * M;
    const uint32_t nch = (it.end - it.begin + k3NK - 1) / k3NK;
    // sample loads: lane = (entry kk, coil quarter cq), 2 coils each
    constexpr int kLC = k3Coils / 4;
    const int kk = lane & (k3NK - 1), cq = lane >> 4;
    const float2* s2 = reinterpret_cast<const float2*>(samples);
    // gridding: lane = (plane zp, coil, re/im)
    const int zp = lane >> 4, q16 = lane & 15;

    float acc[k3T0 * k3T1];
#pragma unroll
    for (int i = 0; i < k3T0 * k3T1; i++) acc[i] = 0.0f;

    constexpr int kDV = DM == 2 ? kLC : 1;
    float4 w0, w1;
    uint32_t hreg;
    float2 sv[kLC];
    float dv[kDV];
#pragma unroll
    for (int u = 0; u < kLC; u++) sv[u] = make_float2(0.0f, 0.0f);
#pragma unroll
    for (int u = 0; u < kDV; u++) dv[u] = 1.0f;

    auto stage = [&](int b) {
        float4* d = reinterpret_cast<float4*>(sw[b]);
        d[lane] = w0;
        if (lane < k3NK * k3NW / 4 - kLanes) d[kLanes + lane] = w1;
        if (lane < 2 * k3NK) sh[b][lane] = hreg;
    };
    auto issue = [&](int b) {
        const uint32_t j = sh[b][2 * kk];
#pragma unroll
        for (int u = 0; u < kLC; u++) {
            int c = c0 + cq * kLC + u;
            c = c < C ? c : c0;  // coils past the last one repeat a valid coil, never stored
            sv[u] = s2[(long long)c * FM + j];
            if constexpr (DM == 2) dv[u] = dcw[dcw_index(c, j, FM, dcw_len)];
        }
        if constexpr (DM == 1) dv[0] = dcw[j];
    };
    auto transpose = [&]() {
#pragma unroll
        for (int u = 0; u < kLC; u++) {
            float2 v = sv[u];
            if constexpr (DM != 0) {
                const float d = dv[DM == 2 ? u : 0];
                v.x *= d;
                v.y *= d;
            }
            *reinterpret_cast<float2*>(&xs[kk * k3XsRow + 2 * (cq * kLC + u)]) = v;
        }
    };
    auto grid_chunk = [&](int b) {
        const int hv = (int)sh[b][lane & 31];
        const float* wb = sw[b];
#pragma unroll
        for (int k = 0; k < k3NK; k++) {
            const float4 x0 = *reinterpret_cast<const float4*>(wb + k * k3NW);
            const float4 x1 = *reinterpret_cast<const float4*>(wb + k * k3NW + 4);
            const float4 y0 = *reinterpret_cast<const float4*>(wb + k * k3NW + 8);
            const float4 y1 = *reinterpret_cast<const float4*>(wb + k * k3NW + 12);
            const float wz = wb[k * k3NW + 16 + zp];
            const float s = xs[k * k3XsRow + q16] * wz;  // dcw already applied in the transpose
            const uint32_t rows = (uint32_t)__builtin_amdgcn_readlane(hv, 2 * k + 1);
            const float wc[k3T0] = {x0.x, x0.y, x0.z, x0.w, x1.x, x1.y, x1.z, x1.w};
            const float wr[k3T1] = {y0.x, y0.y, y0.z, y0.w, y1.x, y1.y, y1.z, y1.w};
#pragma unroll
            for (int cy = 0; cy < k3T1; cy++) {
                if (!((rows >> cy) & 1u)) continue;
                const float sy = s * wr[cy];
#pragma unroll
                for (int cx = 0; cx < k3T0; cx++) acc[cy * k3T0 + cx] = fmaf(sy, wc[cx], acc[cy * k3T0 + cx]);
            }
        }
    };

    if (nch > 0) {
        nc2c3_chunk_load(wt, hd, it.begin, it.end, lane, w0, w1, hreg);
        stage(0);
        nc2c3_chunk_load(wt, hd, it.begin + k3NK, it.end, lane, w0, w1, hreg);
        issue(0);
        for (uint32_t c = 0; c < nch; c++) {
            const int b = (int)(c & 1u);
            transpose();
            stage(b ^ 1);
            nc2c3_chunk_load(wt, hd, it.begin + (c + 2) * k3NK, it.end, lane, w0, w1, hreg);
            issue(b ^ 1);
            grid_chunk(b);
        }
    }

    // write-out through LDS, half a plane tile (4 rows of 8 cells) per pass:
    // lanes run along x within (plane, coil, y) rows for coalesced stores
    const int f = (int)(it.key / (uint32_t)g.ntiles_total);
    int org[3];
    tile_origin3((int)(it.key % (uint32_t)g.ntiles_total), g, org);
    const int lr = lane >> 4, li = lane & 15;
    const int lx = li >> 1, comp = li & 1;
    const int gx = org[0] + lx;
#pragma unroll
    for (int half = 0; half < 2; half++) {
        __syncthreads();
#pragma unroll
        for (int c = 0; c < 32; c++) xp[c * kSubP + lane] = acc[half * 32 + c];
        __syncthreads();
        // rows r = (plane, coil, y in half): 4 x 8 x 4 = 128, four per pass
        for (int r = lr; r < k3T2 * k3Coils * 4; r += 4) {
            const int ly4 = r & 3, cl = (r >> 2) & (k3Coils - 1), pz = r >> 5;
            const int ly = half * 4 + ly4;
            const int cc = c0 + cl;
            const int gy = org[1] + ly, gz = org[2] + pz;
            if (cc >= C || gx >= g.nos[0] || gy >= g.nos[1] || gz >= g.nos[2]) continue;
            float v = xp[(ly4 * k3T0 + lx) * kSubP + pz * 16 + 2 * cl + comp];
            if (it.partial >= 0) {
                const int cell = (pz * k3T1 + ly) * k3T0 + lx;
                partials[(((long long)it.partial * C + cc) * (k3T0 * k3T1 * k3T2) + cell) * 2 + comp] = v;
                continue;
            }
            if (sign && ((gx + gy + gz) & 1)) v = -v;
            const long long o =
                2 * ((long long)gx + L.sy * gy + L.sz * gz + L.sb * ((long long)cc * F + f)) + comp;
            if (accumulate) v += grid[o];
            grid[o] = v;
        }
    }
}

// ---------------------------------------------------------------------------
// nc2c_mfma3d_kernel: the 3D NC2C on the matrix cores.  One wave owns one
// 8 x 8 x 4 tile for 8 coils, 16 rows q: the real parts of the 8 coils, then
// their imaginary parts.  The tile is a product grid[q][cell] = sum over
// entries k of s[q][k] w[k][cell], w[k][cell] = wx_k[x] wy_k[y] wz_k[z]: each
// chunk of 32 table entries is one K = 32 step of v_mfma_f32_16x16x32_f16 on
// 16 blocks of 16 cells (block = plane z and rows 2p, 2p + 1), with the
// split-fp16 products hi*hi + hi*lo + lo*hi and the power-of-two scales of
// nc2c_mfma2d_kernel (relative error about 2^-22 per product, FP32 sums).  Per
// chunk the products wy_k[y] wz_k[z] of all 32 (y, z) pairs are formed once in
// LDS, so a block's B operand is 8 products with the lane's wx_k[x].  The
// table, its chunks and the items are those of nc2c_reg3d_kernel, staged in
// LDS transposed (entries contiguous); entry and product order are fixed, so
// the result is bitwise repeatable.
// Operand lane layouts (16x16x32 f16): A holds row l % 16 at k = 8 (l / 16) +
// t, B column l % 16 at the same k; accumulator register r of lane l is row
// 4 (l / 16) + r, column l % 16.
// ---------------------------------------------------------------------------
constexpr int k3MK = 32;         // table entries per chunk (the MFMA's K)
constexpr int k3MS = k3MK + 4;   // floats per LDS row of entry-contiguous data
struct Mf3Lds {
    static constexpr int Tab = 2 * k3NW * k3MS;     // two chunks' weights [chunk][component][entry]
    static constexpr int Heads = 2 * 2 * k3MK;      // two chunks' (j, rows) heads
    static constexpr int Xs = 16 * k3MS;            // samples [row q][entry]
    static constexpr int Wyz = k3T1 * k3T2 * k3MS;  // wy wz products [y * 4 + z][entry]
    static constexpr int Loop = Tab + Heads + Xs + Wyz;
    static constexpr int Out = 16 * (k3T0 * k3T1 + 1);  // write-out of one plane [q][cell]
    static constexpr int Floats = ((Out > Loop ? Out : Loop) + 3) & ~3;
};
static_assert(k3NW % 4 == 0 && k3MK * k3NW / 4 <= 3 * kLanes, "a chunk's weights load as three float4 per lane");

template <int DM>
__global__ __launch_bounds__(64, 3) void nc2c_mfma3d_kernel(
    const float* __restrict__ samples, float* __restrict__ grid, float* __restrict__ partials,
    const float* __restrict__ wt, const uint32_t* __restrict__ hd, const WorkItem* __restrict__ items, int M,
    int F, int C, GridGeom g, Layout L, int sign, const float* __restrict__ dcw, long long dcw_len,
    int accumulate, float wscale) {
    constexpr int kF4 = k3NW / 4;  // float4 per table entry
    constexpr int kOut = k3T0 * k3T1 + 1;
    __shared__ float4 smem[Mf3Lds::Floats / 4];
    float* tT = reinterpret_cast<float*>(smem);                     // [2][k3NW][k3MS]
    uint32_t* sh = reinterpret_cast<uint32_t*>(tT + Mf3Lds::Tab);   // [2][2 k3MK]
    float* xs = reinterpret_cast<float*>(sh + Mf3Lds::Heads);       // [16][k3MS]
    float* wyz = xs + Mf3Lds::Xs;                                   // [32][k3MS]
    float* xp = reinterpret_cast<float*>(smem);                     // write-out [16][kOut]
    const int lane = threadIdx.x;
    const WorkItem it = items[blockIdx.x];
    if (it.partial == kNoopItem) return;
    const int c0 = blockIdx.y * k3Coils;
    const long long FM = (long long)F * M;
    const uint32_t nch = (it.end - it.begin + k3MK - 1) / k3MK;
    const int row = lane & 15, kh = lane >> 4;            // MFMA: row / column l % 16, entries 8 kh ..
    const int cx = lane & 7, cyo = (lane & 15) >> 3;      // the lane's B column: x, y offset in a block
    const int se = lane & 31, ch = lane >> 5;             // sample loads: entry, coil half (4 coils)
    const float2* s2 = reinterpret_cast<const float2*>(samples);

    mf_float4 acc[16];  // block = z * 4 + p
#pragma unroll
    for (int b = 0; b < 16; b++) acc[b] = mf_float4{0.0f, 0.0f, 0.0f, 0.0f};
    float R = 0.0f;  // sample scale (0: not set yet)

    constexpr int kDV = DM == 2 ? 4 : 1;
    float4 wreg[3];
    uint32_t hreg = 0u;
    float2 sv[4];
    float dv[kDV];
#pragma unroll
    for (int u = 0; u < 4; u++) sv[u] = make_float2(0.0f, 0.0f);
#pragma unroll
    for (int u = 0; u < kDV; u++) dv[u] = 1.0f;

    // entries [cb, cb + 32) of [begin, end), end > begin; entries past end
    // read as sample 0 with zero weights (clamped addresses)
    auto load = [&](uint32_t cb) {
        const float4* w4 = reinterpret_cast<const float4*>(wt);
#pragma unroll
        for (int t = 0; t < 3; t++) {
            const int f = lane + kLanes * t;
            const uint32_t e = cb + (uint32_t)(f / kF4);
            const float4 v = w4[(size_t)min(e, it.end - 1) * kF4 + f % kF4];
            wreg[t] = (f < k3MK * kF4 && e < it.end) ? v : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        }
        const uint32_t eh = cb + (uint32_t)(lane >> 1);
        const uint32_t hv = hd[(size_t)min(eh, it.end - 1) * 2 + (lane & 1)];
        hreg = eh < it.end ? hv : 0u;
    };
    auto stage = [&](int b) {
#pragma unroll
        for (int t = 0; t < 3; t++) {
            const int f = lane + kLanes * t;
            if (f < k3MK * kF4) {
                const int e = f / kF4, c = 4 * (f % kF4);
                float* d = tT + (b * k3NW + c) * k3MS + e;
                d[0] = wreg[t].x;
                d[k3MS] = wreg[t].y;
                d[2 * k3MS] = wreg[t].z;
                d[3 * k3MS] = wreg[t].w;
            }
        }
        sh[b * 2 * k3MK + lane] = hreg;
    };
    auto issue = [&](int b) {
        const uint32_t j = sh[b * 2 * k3MK + 2 * se];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            int c = c0 + 4 * ch + u;
            c = c < C ? c : c0;  // coils past the last one repeat a valid coil, never stored
            sv[u] = s2[(long long)c * FM + j];
            if constexpr (DM == 2) dv[u] = dcw[dcw_index(c, j, FM, dcw_len)];
        }
        if constexpr (DM == 1) dv[0] = dcw[j];
    };
    auto transpose = [&]() {
#pragma unroll
        for (int u = 0; u < 4; u++) {
            float2 v = sv[u];
            if constexpr (DM != 0) {
                const float d = dv[DM == 2 ? u : 0];
                v.x *= d;
                v.y *= d;
            }
            const int cl = 4 * ch + u;
            xs[cl * k3MS + se] = v.x;
            xs[(k3Coils + cl) * k3MS + se] = v.y;
        }
    };
    auto grid_chunk = [&](int b) {
        const float* tb = tT + b * k3NW * k3MS;
        {
            // wy_k[y] wz_k[z] for pair (y, z) = lane / 2, entries 16 (lane & 1) ..
            const int pr = lane >> 1, y = pr >> 2, z = pr & 3, k0 = 16 * (lane & 1);
            const float4* py = reinterpret_cast<const float4*>(tb + (k3T0 + y) * k3MS + k0);
            const float4* pz = reinterpret_cast<const float4*>(tb + (k3T0 + k3T1 + z) * k3MS + k0);
            float4* d = reinterpret_cast<float4*>(wyz + pr * k3MS + k0);
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const float4 a = py[q], c = pz[q];
                d[q] = make_float4(a.x * c.x, a.y * c.y, a.z * c.z, a.w * c.w);
            }
        }
        __syncthreads();
        // A: row `row`, entries 8 kh .. 8 kh + 7
        const float Rt = R == 0.0f ? 1.0f : R;
        const float4* xa = reinterpret_cast<const float4*>(xs + row * k3MS + 8 * kh);
        const float4 a0 = xa[0], a1 = xa[1];
        mf_float8 x = mf_float8{a0.x, a0.y, a0.z, a0.w, a1.x, a1.y, a1.z, a1.w} * Rt;
        const float m = mf_absmax(x);
        if (__builtin_amdgcn_ballot_w64(m > 32768.0f || (R == 0.0f && m > 0.0f))) {
            const float mx = mf_wave_max(m);  // of x = samples * Rt
            int e;
            frexpf(mx, &e);                   // mx < 2^e
            e = min(max(e, -100), 140);       // keep R and 1 / R finite
            const float fs = ldexpf(1.0f, 12 - e);
            if (R != 0.0f) {
#pragma unroll
                for (int bl = 0; bl < 16; bl++) acc[bl] *= fs;  // powers of two: exact
            }
            x *= fs;
            R = Rt * fs;
        }
        mf_half8 xh, xl;
        mf_split(x, xh, xl);
        // B: the lane's x weights for its 8 entries (scaled), times wy wz of each block
        const float4* px = reinterpret_cast<const float4*>(tb + cx * k3MS + 8 * kh);
        const float4 x0 = px[0], x1 = px[1];
        const mf_float8 wx = mf_float8{x0.x, x0.y, x0.z, x0.w, x1.x, x1.y, x1.z, x1.w} * wscale;
#pragma unroll
        for (int z = 0; z < k3T2; z++) {
#pragma unroll
            for (int pp = 0; pp < 4; pp += 2) {
                // two blocks at a time: the accumulators alternate between their products
                mf_half8 bh[2], bl[2];
#pragma unroll
                for (int h = 0; h < 2; h++) {
                    const int y = 2 * (pp + h) + cyo;
                    const float4* pw = reinterpret_cast<const float4*>(wyz + (y * k3T2 + z) * k3MS + 8 * kh);
                    const float4 w0 = pw[0], w1 = pw[1];
                    mf_split(wx * mf_float8{w0.x, w0.y, w0.z, w0.w, w1.x, w1.y, w1.z, w1.w}, bh[h], bl[h]);
                }
                const int b0 = z * 4 + pp, b1 = b0 + 1;
                acc[b0] = __builtin_amdgcn_mfma_f32_16x16x32_f16(xl, bh[0], acc[b0], 0, 0, 0);
                acc[b1] = __builtin_amdgcn_mfma_f32_16x16x32_f16(xl, bh[1], acc[b1], 0, 0, 0);
                acc[b0] = __builtin_amdgcn_mfma_f32_16x16x32_f16(xh, bl[0], acc[b0], 0, 0, 0);
                acc[b1] = __builtin_amdgcn_mfma_f32_16x16x32_f16(xh, bl[1], acc[b1], 0, 0, 0);
                acc[b0] = __builtin_amdgcn_mfma_f32_16x16x32_f16(xh, bh[0], acc[b0], 0, 0, 0);
                acc[b1] = __builtin_amdgcn_mfma_f32_16x16x32_f16(xh, bh[1], acc[b1], 0, 0, 0);
            }
        }
    };

    if (nch > 0) {
        load(it.begin);
        stage(0);
        load(it.begin + k3MK);
        __syncthreads();
        issue(0);
        for (uint32_t c = 0; c < nch; c++) {
            const int b = (int)(c & 1u);
            transpose();
            stage(b ^ 1);
            load(it.begin + (c + 2) * k3MK);
            __syncthreads();
            issue(b ^ 1);
            grid_chunk(b);
            __syncthreads();
        }
    }

    // write-out through LDS one plane at a time: xp[q][y * 8 + x], then lanes
    // run along (re/im, x) for coalesced stores of (coil, y) rows
    const float outs = R > 0.0f ? 1.0f / (R * wscale) : 0.0f;  // powers of two: exact
    const int f = (int)(it.key / (uint32_t)g.ntiles_total);
    int org[3];
    tile_origin3((int)(it.key % (uint32_t)g.ntiles_total), g, org);
#pragma unroll
    for (int z = 0; z < k3T2; z++) {
        __syncthreads();
#pragma unroll
        for (int pp = 0; pp < 4; pp++)
#pragma unroll
            for (int r = 0; r < 4; r++)
                xp[(4 * kh + r) * kOut + (2 * pp + cyo) * k3T0 + cx] = acc[z * 4 + pp][r] * outs;
        __syncthreads();
        const int gz = org[2] + z;
#pragma unroll 4
        for (int t = 0; t < 16; t++) {
            const int idx = lane + kLanes * t;  // 8 coils x 8 rows x 8 columns x (re, im)
            const int comp = idx & 1, lx = (idx >> 1) & 7, ly = (idx >> 4) & 7, cl = idx >> 7;
            const int cc = c0 + cl, gx = org[0] + lx, gy = org[1] + ly;
            if (cc >= C || gx >= g.nos[0] || gy >= g.nos[1] || gz >= g.nos[2]) continue;
            float v = xp[(comp * k3Coils + cl) * kOut + ly * k3T0 + lx];
            if (it.partial >= 0) {
                const int cell = (z * k3T1 + ly) * k3T0 + lx;
                partials[(((long long)it.partial * C + cc) * (k3T0 * k3T1 * k3T2) + cell) * 2 + comp] = v;
                continue;
            }
            if (sign && ((gx + gy + gz) & 1)) v = -v;
            const long long o =
                2 * ((long long)gx + L.sy * gy + L.sz * gz + L.sb * ((long long)cc * F + f)) + comp;
            if (accumulate) v += grid[o];
            grid[o] = v;
        }
    }
}

// Sum of the partial tiles of split tiles in segment order; one thread per
// (coil, cell, re/im) element of a tile.
__global__ __launch_bounds__(256) void nc2c_reduce3d_kernel(const float* __restrict__ partials,
                                                            float* __restrict__ grid,
                                                            const RedItem* __restrict__ red, int F, int C,
                                                            GridGeom g, Layout L, int sign, int accumulate) {
    constexpr int kCells = k3T0 * k3T1 * k3T2;
    constexpr int kTileFloats = kCells * 2;
    const RedItem it = red[blockIdx.x];
    const int i = blockIdx.y * blockDim.x + threadIdx.x;
    if (i >= C * kTileFloats) return;
    const int comp = i & 1, cell = (i >> 1) % kCells, cc = i / kTileFloats;
    const int f = (int)(it.key / (uint32_t)g.ntiles_total);
    int org[3];
    tile_origin3((int)(it.key % (uint32_t)g.ntiles_total), g, org);
    const int gx = org[0] + cell % k3T0;
    const int gy = org[1] + (cell / k3T0) % k3T1;
    const int gz = org[2] + cell / (k3T0 * k3T1);
    if (gx >= g.nos[0] || gy >= g.nos[1] || gz >= g.nos[2]) return;
    const long long stride = (long long)C * kTileFloats;
    const float* src = partials + (long long)it.first * stride + i;
    float a0 = 0.0f, a1 = 0.0f, a2 = 0.0f, a3 = 0.0f;
    int s = 0;
    for (; s + 4 <= it.count; s += 4) {
        a0 += src[(s + 0) * stride];
        a1 += src[(s + 1) * stride];
        a2 += src[(s + 2) * stride];
        a3 += src[(s + 3) * stride];
    }
    for (; s < it.count; s++) a0 += src[s * stride];
    float v = (a0 + a1) + (a2 + a3);
    if (sign && ((gx + gy + gz) & 1)) v = -v;
    const long long o = 2 * ((long long)gx + L.sy * gy + L.sz * gz + L.sb * ((long long)cc * F + f)) + comp;
    if (accumulate) v += grid[o];
    grid[o] = v;
}

// ---- C2NC ------------------------------------------------------------------
constexpr int k3S0 = 4, k3S1 = 4, k3S2 = 4;  // staged tile (x, y, z), plus the kernel halo
constexpr int k3CDw = 24;                    // dwords per C2NC table entry
constexpr int k3XSlots = 8;                  // x weight slots (two halves of up to 4)
constexpr int k3Chunk = 64;                  // C2NC table entries staged in LDS at a time
constexpr int k3Seg = 512;                   // entries per C2NC work item
constexpr int k3Waves = 8;                   // wavefronts per C2NC workgroup
constexpr int k3Cell = 2 * k3Coils;          // LDS floats per staged cell

// C2NC table entry (24 dwords):
//   [0] sample index j
//   [1] bits 0-7 bx, 8-15 by, 16-23 bz: window slot 0 in the staged tile
//   [2..9] x weights of slots 0..7, [10..16] y weights of slots 0..6,
//   [17..23] z weights of slots 0..6 (zero past the window)
template <int KW>
__global__ void c2nc_table3d_kernel(const float* __restrict__ p, const uint32_t* __restrict__ list,
                                    long long n, GridGeom g, uint32_t* __restrict__ tab) {
    const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const uint32_t j = list[e];
    const int fr = (int)floorf(g.radius);
    uint32_t* t = tab + e * k3CDw;
    uint32_t packed = 0;
#pragma unroll
    for (int d = 0; d < 3; d++) {
        float pw;
        const int td = home_tile_1d(p[(long long)j * 3 + d], g, d, pw);
        float w[KW];
        int lo;
        window_raw<KW>(pw, g, d, w, lo);
        // 0 <= b <= tile for every home sample; the clamp keeps reads inside LDS regardless
        const int b = min(max(lo - (td * g.tile[d] - fr), 0), g.tile[d]);
        packed |= (uint32_t)b << (8 * d);
        const int base = d == 0 ? 2 : (d == 1 ? 10 : 17);
        const int ns = d == 0 ? k3XSlots : kTabKW;
#pragma unroll
        for (int i = 0; i < ns; i++) t[base + i] = __float_as_uint(i < KW ? w[i < KW ? i : 0] : 0.0f);
    }
    t[0] = j;
    t[1] = packed;
}

// LDS of the C2NC kernel: one plane per coil holding the staged cells
// E0 x E1 x E2 (c2nc_extent per dimension) as (re, im) pairs in the grid's
// own order (x fastest), then zeroed pad cells -- a window reads up to one cell,
// row and plane past the staged tile (and, with the x halves, 2 XH - KW more
// cells), always with weight 0.  The plane stride is 4 (mod 32) floats: the 8
// coil lanes of one x half, reading (re, im), then fall on banks {4c, 4c + 1},
// and those of the other half, 2 XH floats on (XH = 3 for W 5.5), on the
// remaining 16 banks.
__host__ __device__ constexpr int c2nc3_cells(int KW) {
    return c2nc_extent(k3S0, KW) * c2nc_extent(k3S1, KW) * c2nc_extent(k3S2, KW);
}
__host__ __device__ constexpr int c2nc3_pad_cells(int KW) {
    return c2nc_extent(k3S0, KW) * c2nc_extent(k3S1, KW) + c2nc_extent(k3S0, KW) + 8;
}
__host__ __device__ constexpr int c2nc3_plane_floats(int KW) {
    return ((2 * (c2nc3_cells(KW) + c2nc3_pad_cells(KW)) + 31) / 32) * 32 + 4;
}
__host__ __device__ constexpr int c2nc3_lds_floats(int KW) { return k3Coils * c2nc3_plane_floats(KW); }
__host__ __device__ constexpr int c2nc3_lds_bytes(int KW) { return (c2nc3_lds_floats(KW) + k3Chunk * k3CDw + k3Chunk) * 4; }

// Staging of one work item into registers: its tile + halo for 8 coils, and
// one chunk (k3Chunk entries) of its table.  Wave w stages coil c0 + w; its
// lanes (row r, float f) cover R grid rows of 2 E0 floats at a time, so a load
// instruction reads R runs of 2 E0 consecutive floats and the LDS stores are
// contiguous.  Row and plane wraps and centring signs are worked out once per
// row and plane.
template <int KW>
struct C2NC3Stage {
    static constexpr int E0 = c2nc_extent(k3S0, KW), E1 = c2nc_extent(k3S1, KW), E2 = c2nc_extent(k3S2, KW);
    static constexpr int NT = 64 * k3Waves;
    static constexpr int RF = 2 * E0;               // floats per staged row
    static constexpr int R = kLanes / RF;           // rows per load instruction
    static constexpr int kYI = (E1 + R - 1) / R;    // load instructions per plane
    static constexpr int kU4 = k3CDw / 4;           // table uint4 per entry
    static_assert(k3Chunk * kU4 <= NT, "one table uint4 per thread");
    static_assert(k3Waves == k3Coils, "one wave per coil");
    float v[E2][kYI];
    int par;  // bit m * E2 + ez: centring sign of value (ez, m)
    uint4 ta;
    float dj;  // dcw of chunk entry tid (per-entry dcw array, or 1)

    __device__ __forceinline__ void load_grid(const float* __restrict__ grid, uint32_t key, bool live, int tid,
                                              int c0, int C, int F, const GridGeom& g, const Layout& L, int sign) {
        const int fr = (int)floorf(g.radius);
        const int f = (int)(key / (uint32_t)g.ntiles_total);
        int org[3];
        tile_origin3((int)(key % (uint32_t)g.ntiles_total), g, org);
        const int sx = org[0] - fr, sy = org[1] - fr, sz = org[2] - fr;
        const int lane = tid & 63, coil = c0 + (tid >> 6);
        const int r = lane / RF, fl = lane % RF;  // row in the instruction, float in the row
        const int ex = fl >> 1, comp = fl & 1;
        const bool cok = live && r < R && coil < C;
        int gx = sx + ex;
        const int px = gx & 1;
        gx = gx < 0 ? gx + g.nos[0] : (gx >= g.nos[0] ? gx - g.nos[0] : gx);
        const float* gb = grid + 2 * ((long long)gx + L.sb * ((long long)(cok ? coil : c0) * F + f)) + comp;
        long long oy[kYI];
        int py[kYI];
#pragma unroll
        for (int m = 0; m < kYI; m++) {
            int gy = sy + r + m * R;
            py[m] = gy & 1;
            gy = gy < 0 ? gy + g.nos[1] : (gy >= g.nos[1] ? gy - g.nos[1] : gy);
            oy[m] = 2 * L.sy * gy;
        }
        par = 0;
#pragma unroll
        for (int ez = 0; ez < E2; ez++) {
            int gz = sz + ez;
            // the parity of a wrapped coordinate equals that of the unwrapped one (even N_os)
            const int pz = gz & 1;
            gz = gz < 0 ? gz + g.nos[2] : (gz >= g.nos[2] ? gz - g.nos[2] : gz);
            const float* pl = gb + 2 * L.sz * gz;
#pragma unroll
            for (int m = 0; m < kYI; m++) {
                const bool ok = cok && r + m * R < E1;
                v[ez][m] = ok ? pl[oy[m]] : 0.0f;
                par |= (sign && ((px ^ py[m] ^ pz) & 1)) ? (1 << (m * E2 + ez)) : 0;
            }
        }
    }

    // entries [e0, e0 + n) of the table, n <= k3Chunk
    __device__ __forceinline__ void load_table(const uint32_t* __restrict__ tab, const float* __restrict__ dcw_e,
                                               uint32_t e0, int n, int tid) {
        ta = make_uint4(0u, 0u, 0u, 0u);
        if (tid < n * kU4) ta = reinterpret_cast<const uint4*>(tab)[(size_t)e0 * kU4 + tid];
        dj = 1.0f;
        if (dcw_e && tid < n) dj = dcw_e[e0 + tid];
    }

    __device__ __forceinline__ void store_grid(float* sg, int tid) const {
        const int lane = tid & 63;
        const int r = lane / RF, fl = lane % RF;
        if (r < R) {
            float* d = sg + (tid >> 6) * c2nc3_plane_floats(KW) + r * RF + fl;
#pragma unroll
            for (int ez = 0; ez < E2; ez++)
#pragma unroll
                for (int m = 0; m < kYI; m++) {
                    if (r + m * R >= E1) continue;
                    const float x = v[ez][m];
                    d[(ez * E1 + m * R) * RF] = ((par >> (m * E2 + ez)) & 1) ? -x : x;
                }
        }
    }

    __device__ __forceinline__ void store_table(uint32_t* st, float* sd, int n, int tid) const {
        if (tid < n * kU4) reinterpret_cast<uint4*>(st)[tid] = ta;
        if (tid < n) sd[tid] = dj;
    }
};

// dcw of every C2NC table entry (the samples' dcw in table order), so that the
// gridding kernel loads it with the table instead of through a dependent load.
__global__ void entry_dcw_kernel(const uint32_t* __restrict__ list, long long n, const float* __restrict__ dcw,
                                 float* __restrict__ out) {
    const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < n) out[e] = dcw[list[e]];
}

// MODE 0: no dcw, no accumulation; 1: one dcw weight per sample (dcw_e: per
// table entry), no accumulation; 2: anything else.
// Persistent: workgroup b takes items b, b + gridDim.x, ... (gridDim.x a
// multiple of 8 keeps each item on its XCD).  An item stages its tile once and
// walks its table in chunks of k3Chunk entries; the next chunk, or after the
// last one the next item's staging, is in flight while a chunk is gathered.
template <int KW, int MODE>
__global__ __launch_bounds__(64 * k3Waves, 4) void c2nc_multi3d_kernel(
    const float* __restrict__ grid, float* __restrict__ out, const uint32_t* __restrict__ tab,
    const WorkItem* __restrict__ items, int n_items, int M, int F, int C, GridGeom g, Layout L, int sign,
    const float* __restrict__ dcw, long long dcw_len, const float* __restrict__ dcw_e, int accumulate) {
    extern __shared__ float sg[];  // coil planes, then a table chunk, then its dcw
    using S = C2NC3Stage<KW>;
    constexpr int E0 = S::E0, E1 = S::E1;
    constexpr int NCELL = c2nc3_cells(KW);
    constexpr int PS = c2nc3_plane_floats(KW);
    constexpr int XH = (KW + 1) / 2;  // x slots per half
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = wave_uniform(tid >> 6);
    const int c0 = blockIdx.y * k3Coils;
    uint32_t* st = reinterpret_cast<uint32_t*>(sg + c2nc3_lds_floats(KW));
    float* sd = reinterpret_cast<float*>(st + k3Chunk * k3CDw);
    const long long FM = (long long)F * M;

    for (int i = tid; i < k3Coils * (PS - 2 * NCELL); i += S::NT) {
        const int c = i / (PS - 2 * NCELL), o = i % (PS - 2 * NCELL);
        sg[c * PS + 2 * NCELL + o] = 0.0f;
    }

    // lane = (slot, x half, coil): a lane reads whole cells (re, im); lanes past
    // the last coil mirror coil c0 (same values, same output address)
    const int slot = lane >> 4, xh = (lane >> 3) & 1, cl = lane & 7;
    const int coil = c0 + cl;
    const bool on = coil < C;
    const int lcoil = on ? coil : c0;
    float2* ol = reinterpret_cast<float2*>(out) + (long long)lcoil * FM;
    const float2* sl = reinterpret_cast<const float2*>(sg + (on ? cl : 0) * PS) + xh * XH;

    const float* de = MODE == 1 ? dcw_e : nullptr;
    auto entries = [](const WorkItem& w) -> uint32_t { return w.partial == kNoopItem ? 0u : w.end - w.begin; };
    auto chunk_n = [](uint32_t ne, uint32_t ch) -> int {
        return (int)min((uint32_t)k3Chunk, ne - min(ne, ch * (uint32_t)k3Chunk));
    };
    S cur;
    int item = blockIdx.x;
    if (item < n_items) {
        const WorkItem it = items[item];
        const uint32_t ne = entries(it);
        cur.load_grid(grid, it.key, ne > 0, tid, c0, C, F, g, L, sign);
        cur.load_table(tab, de, it.begin, chunk_n(ne, 0), tid);
    }
    for (; item < n_items; item += gridDim.x) {
        const WorkItem it = items[item];
        const uint32_t ne = entries(it);
        const uint32_t nch = (ne + k3Chunk - 1) / k3Chunk;
        const int nxt = item + (int)gridDim.x;
        // issues the loads of chunk ch of this item, or (past its last chunk)
        // the next item's staging and first chunk
        auto issue = [&](uint32_t ch) {
            if (ch < nch) {
                cur.load_table(tab, de, it.begin + ch * k3Chunk, chunk_n(ne, ch), tid);
            } else if (nxt < n_items) {
                const WorkItem in = items[nxt];
                const uint32_t nn = entries(in);
                cur.load_grid(grid, in.key, nn > 0, tid, c0, C, F, g, L, sign);
                cur.load_table(tab, de, in.begin, chunk_n(nn, 0), tid);
            }
        };
        __syncthreads();  // the previous item's gather is done with LDS
        cur.store_grid(sg, tid);
        cur.store_table(st, sd, chunk_n(ne, 0), tid);
        __syncthreads();
        issue(1);
        for (uint32_t ch = 0; ch < nch; ch++) {
            const uint32_t n = (uint32_t)chunk_n(ne, ch);
            // wave w takes a contiguous share of the chunk, four entries at a time
            const uint32_t q0 = (n * (uint32_t)wave) / k3Waves, q1 = (n * (uint32_t)(wave + 1)) / k3Waves;
            for (uint32_t k = q0; k < q1; k += 4) {
                const uint32_t e = min(k + (uint32_t)slot, q1 - 1);  // a short tail repeats the last entry
                const uint32_t* te = st + e * k3CDw;
                const uint32_t j = te[0];
                const uint32_t pk = te[1];
                const int bx = (int)(pk & 0xff), by = (int)((pk >> 8) & 0xff), bz = (int)((pk >> 16) & 0xff);
                float wx[XH], wy[KW], wz[KW];
#pragma unroll
                for (int i = 0; i < XH; i++) wx[i] = __uint_as_float(te[2 + xh * XH + i]);
#pragma unroll
                for (int i = 0; i < KW; i++) {
                    wy[i] = __uint_as_float(te[10 + i]);
                    wz[i] = __uint_as_float(te[17 + i]);
                }
                const float2* base = sl + ((bz * E1 + by) * E0 + bx);
                float ar = 0.0f, ai = 0.0f;
#pragma unroll
                for (int iz = 0; iz < KW; iz++) {
                    float2 v[KW][XH];
#pragma unroll
                    for (int iy = 0; iy < KW; iy++)
#pragma unroll
                        for (int ix = 0; ix < XH; ix++) v[iy][ix] = base[(iz * E1 + iy) * E0 + ix];
                    float pr = 0.0f, pi = 0.0f;
#pragma unroll
                    for (int iy = 0; iy < KW; iy++) {
                        float rr = 0.0f, ri = 0.0f;
#pragma unroll
                        for (int ix = 0; ix < XH; ix++) {
                            rr += wx[ix] * v[iy][ix].x;
                            ri += wx[ix] * v[iy][ix].y;
                        }
                        pr += wy[iy] * rr;
                        pi += wy[iy] * ri;
                    }
                    ar += wz[iz] * pr;
                    ai += wz[iz] * pi;
                }
                // add the other x half (lane ^ 8)
                ar += __int_as_float(__builtin_amdgcn_ds_swizzle(__float_as_int(ar), 0x201F));
                ai += __int_as_float(__builtin_amdgcn_ds_swizzle(__float_as_int(ai), 0x201F));
                if (xh == 0) {
                    float d = 1.0f;
                    if constexpr (MODE == 1) d = sd[e];
                    if constexpr (MODE == 2) {
                        const int dm = dcw_mode(dcw, dcw_len, FM);
                        if (dm == 1) d = dcw[j];
                        else if (dm == 2) d = dcw[dcw_index(lcoil, j, FM, dcw_len)];
                    }
                    float2 a = make_float2(ar * d, ai * d);
                    if (MODE == 2 && accumulate) {
                        const float2 old = ol[j];
                        a.x += old.x;
                        a.y += old.y;
                    }
                    ol[j] = a;
                }
            }
            if (ch + 1 < nch) {
                __syncthreads();  // the chunk's gather is done with the table
                cur.store_table(st, sd, chunk_n(ne, ch + 1), tid);
                __syncthreads();
                issue(ch + 2);
            }
        }
    }
}

// ---------------------------------------------------------------------------
// c2nc_mfma3d_kernel: the 3D C2NC of 8 coils per workgroup on the matrix cores.
//
// Samples are binned by home tile of T x Ty x T cells, T = c2nc_mfma3_tile(KW)
// (3 for windows of 5 or 6 cells) and Ty = c2nc_mfma3_tile(KW, EY), so that a
// tile and its kernel halo span 8 staged cells in x and z and EY in y (from
// tile origin - floor(r)).  For each staged plane z the samples of a tile are
// a product,
//   P_z[q][n] = sum over staged (x, y) of grid[q][x, y, z] * wx_n[x] * wy_n[y],
// q = (coil, re/im), n = sample, and sample n of row q is sum_z wz_n[z] P_z[q][n].
// A work item is a run of tiles along z in one column (tx, ty): the workgroup
// keeps the 8 staged planes of the current tile in a ring of 8 LDS plane
// slots, so the next tile loads only its T new planes.  A plane is staged by
// one wave, scaled by its own power of two R_z (largest |value| below 2^12)
// and split into fp16 hi / lo rows; each wave then takes 16 samples at a time
// and forms P_z as EY / 4 K = 32 steps of v_mfma_f32_16x16x32_f16 (lane group
// l / 16 takes staged row y = 4 h + l / 16 of step h, 8 x values per lane)
// with the three products hi*hi + hi*lo + lo*hi of nc2c_mfma2d_kernel.  The B
// operand wx wy is split once per 16 samples and serves all 8 planes; the z
// weights and 1 / R_z are applied in FP32.  Out: s * oscale (oscale undoes the
// table's power-of-two x and y scales), then dcw.  Product order is fixed:
// the result is bitwise repeatable.
// ---------------------------------------------------------------------------
constexpr int k3ME = 8;                             // staged cells in x and z
constexpr int k3MWaves = 4;                         // wavefronts per workgroup
constexpr int k3MSeg = 2048;                        // entries per work item, at most
// Staged rows in y, EY = 8, 12 or 16: home tiles of T x (EY - E0) x T cells
// with T = 8 - E0, E0 = c2nc_extent(0, KW).  Taller tiles stage fewer halo
// cells per sample (8 x EY cells for 3 x (EY - 5) tiles at KW 6) at EY / 8
// times the matrix-core work per sample.  At 128^3 -> 192^3, 8 coils: EY 8
// 1.59 ms, 12 1.35 ms, 16 1.63 ms (its 68 KB of LDS leave two workgroups
// per CU).
template <int EY>
struct C2M3Geom {
    static_assert(EY == 8 || EY == 12 || EY == 16, "staged rows in y");
    static constexpr int Cells = k3ME * EY;      // cells of a staged plane
    static constexpr int AS = k3ME * Cells + 8;  // halves per LDS row: 8 plane slots + 8
    static constexpr int W = 2 * k3ME + EY;      // floats per weight-table entry: x, y, z
    static constexpr int KS = EY / 4;            // K = 32 steps per plane
    static constexpr int Occ = EY <= 12 ? 3 : 2; // workgroups per CU (LDS)
};
__host__ __device__ constexpr int c2nc_mfma3_tile(int KW, int E = k3ME) { return E - c2nc_extent(0, KW); }

// Tile keys of the matrix-core 3D C2NC: tz fastest along a column (tx, ty),
// the columns in bricks of 4 x 4 (x fastest within a brick, bricks x fastest),
// so that the columns staged at one time on an XCD are neighbours in x and y
// and share their halo cells in its L2.  The key space is padded to whole
// bricks (ntiles_total of this geometry counts the padding).
__device__ __forceinline__ uint32_t column_tile3(int tx, int ty, int tz, const GridGeom& g) {
    const int nbx = (g.ntiles[0] + 3) >> 2;
    const int col = ((((ty >> 2) * nbx) + (tx >> 2)) << 4) | ((ty & 3) << 2) | (tx & 3);
    return (uint32_t)col * (uint32_t)g.ntiles[2] + (uint32_t)tz;
}

__device__ __forceinline__ void column_tile_coords3(uint32_t t, const GridGeom& g, int& tx, int& ty, int& tz) {
    const int nbx = (g.ntiles[0] + 3) >> 2;
    tz = (int)(t % (uint32_t)g.ntiles[2]);
    const int col = (int)(t / (uint32_t)g.ntiles[2]);
    const int b = col >> 4;
    tx = ((b % nbx) << 2) | (col & 3);
    ty = ((b / nbx) << 2) | ((col >> 2) & 3);
}

// Home tile key of every sample in column order (the matrix-core 3D C2NC).
__global__ void home_keys_column3_kernel(const float* __restrict__ p, long long nsamp, long long M, GridGeom g,
                                         uint32_t* __restrict__ keys, uint32_t* __restrict__ vals) {
    const long long j = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nsamp) return;
    float pw;
    const int tx = home_tile_1d(p[j * 3 + 0], g, 0, pw);
    const int ty = home_tile_1d(p[j * 3 + 1], g, 1, pw);
    const int tz = home_tile_1d(p[j * 3 + 2], g, 2, pw);
    keys[j] = (uint32_t)(j / M) * (uint32_t)g.ntiles_total + column_tile3(tx, ty, tz, g);
    vals[j] = (uint32_t)j;
}

// Weight table of the matrix-core 3D C2NC, 16 + EY floats per entry: the
// sample's x weights at the 8 staged columns (scaled by sx), y weights at the
// EY staged rows (scaled by sy), z weights at the 8 staged planes; zero
// outside the window.  sx and sy are powers of two that bring each axis's
// largest weight to at most 1, so every x y product fits fp16.
template <int KW, int EY>
__global__ void c2nc_wtable3d_kernel(const float* __restrict__ p, const uint32_t* __restrict__ list, long long n,
                                     GridGeom g, float sx, float sy, float* __restrict__ wtab) {
    const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    static_assert(KW >= 2, "a window of at least 2 cells");
    const uint32_t j = list[e];
    const int fr = (int)floorf(g.radius);
    float r[3][EY];
#pragma unroll
    for (int d = 0; d < 3; d++) {
        float pw;
        const int td = home_tile_1d(p[(long long)j * 3 + d], g, d, pw);
        float w[KW];
        int lo;
        window_raw<KW>(pw, g, d, w, lo);
        // 0 <= b <= T; window slots past the staged cells have weight 0
        const int b = min(max(lo - (td * g.tile[d] - fr), 0), g.tile[d]);
        const float sc = d == 0 ? sx : (d == 1 ? sy : 1.0f);
#pragma unroll
        for (int c = 0; c < EY; c++) r[d][c] = 0.0f;
#pragma unroll
        for (int i = 0; i < KW; i++)
#pragma unroll
            for (int c = 0; c < EY; c++)
                if (c == b + i) r[d][c] = w[i] * sc;
    }
    float4* t = reinterpret_cast<float4*>(wtab + e * C2M3Geom<EY>::W);
    t[0] = make_float4(r[0][0], r[0][1], r[0][2], r[0][3]);
    t[1] = make_float4(r[0][4], r[0][5], r[0][6], r[0][7]);
#pragma unroll
    for (int q = 0; q < EY / 4; q++) t[2 + q] = make_float4(r[1][4 * q], r[1][4 * q + 1], r[1][4 * q + 2], r[1][4 * q + 3]);
    t[2 + EY / 4] = make_float4(r[2][0], r[2][1], r[2][2], r[2][3]);
    t[3 + EY / 4] = make_float4(r[2][4], r[2][5], r[2][6], r[2][7]);
}

// Largest of a non-negative value over the wavefront, in every lane: xor
// swizzles within each 32-lane half, then the two halves through SGPRs (no
// address registers, unlike a shuffle).
__device__ __forceinline__ float wave_max_nonneg(float v) {
    v = fmaxf(v, __int_as_float(__builtin_amdgcn_ds_swizzle(__float_as_int(v), (1 << 10) | 0x1f)));
    v = fmaxf(v, __int_as_float(__builtin_amdgcn_ds_swizzle(__float_as_int(v), (2 << 10) | 0x1f)));
    v = fmaxf(v, __int_as_float(__builtin_amdgcn_ds_swizzle(__float_as_int(v), (4 << 10) | 0x1f)));
    v = fmaxf(v, __int_as_float(__builtin_amdgcn_ds_swizzle(__float_as_int(v), (8 << 10) | 0x1f)));
    v = fmaxf(v, __int_as_float(__builtin_amdgcn_ds_swizzle(__float_as_int(v), (16 << 10) | 0x1f)));
    return fmaxf(__int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 0)),
                 __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 32)));
}

// The column of a work item: frame, first tile, staging origin, and the
// lane's staged column x = l % 8 of coil l / 8 (wrapped, with its parity).
struct C2M3Column {
    const float2* cb;  // the lane's coil and frame, staged column x, plane 0 row 0
    int sy, sz0;       // staging origin in y and z (unwrapped) of the first tile
    int px;            // parity of the lane's unwrapped x
    bool ok;           // the lane's coil exists

    __device__ __forceinline__ void setup(const float* __restrict__ grid, uint32_t key, int lane, int c0, int C,
                                          int F, const GridGeom& g, const Layout& L) {
        const int fr = (int)floorf(g.radius);
        const int f = (int)(key / (uint32_t)g.ntiles_total);
        int tx, ty, tz;
        column_tile_coords3(key % (uint32_t)g.ntiles_total, g, tx, ty, tz);
        const int x = lane & 7, cl = lane >> 3;
        ok = c0 + cl < C;
        // coordinates lie in [-floor(r), N_os + 8): one wrap each
        int gx = tx * g.tile[0] - fr + x;
        px = gx & 1;  // the parity of a wrapped coordinate equals that of the unwrapped one (even N_os)
        gx = gx < 0 ? gx + g.nos[0] : (gx >= g.nos[0] ? gx - g.nos[0] : gx);
        cb = reinterpret_cast<const float2*>(grid) + (L.sb * ((long long)(ok ? c0 + cl : c0) * F + f) + gx);
        sy = ty * g.tile[1] - fr;
        sz0 = tz * g.tile[2] - fr;
    }
};

// One staged plane (unwrapped z coordinate zz) of a column in one wave: lane
// = (x, coil) holds the plane's EY rows y; a load instruction reads 64-byte
// row segments.  The centring signs are kept as bits and applied when the
// plane is split into LDS, so the loads stay in flight until then.
template <int EY>
struct C2M3Plane {
    using G = C2M3Geom<EY>;
    float2 v[EY];
    uint32_t par;  // bit y: centring sign of row y

    __device__ __forceinline__ void load(const C2M3Column& c, int zz, const GridGeom& g, const Layout& L, int sign) {
        const int gz = zz < 0 ? zz + g.nos[2] : (zz >= g.nos[2] ? zz - g.nos[2] : zz);
        const float2* pl = c.cb + L.sz * (long long)gz;
        par = 0;
#pragma unroll
        for (int y = 0; y < EY; y++) {
            int gy = c.sy + y;
            const int p = (c.px + gy + zz) & 1;
            gy = gy < 0 ? gy + g.nos[1] : (gy >= g.nos[1] ? gy - g.nos[1] : gy);
            v[y] = c.ok ? pl[L.sy * (long long)gy] : make_float2(0.0f, 0.0f);
            par |= (uint32_t)(sign && p) << y;
        }
    }

    // scale by the plane's power of two R, centring signs, split into fp16 hi
    // / lo rows 2 coil (re), 2 coil + 1 (im) of LDS plane slot `slot`; 1 / R
    // goes to inv[slot]
    __device__ __forceinline__ void store(_Float16* Ahi, _Float16* Alo, float* inv, int slot, int lane) const {
        float m = 0.0f;
#pragma unroll
        for (int y = 0; y < EY; y++) m = fmaxf(m, fmaxf(fabsf(v[y].x), fabsf(v[y].y)));
        const float mx = wave_max_nonneg(m);
        int e2;
        frexpf(mx, &e2);               // mx < 2^e2, e2 <= 128 for finite values
        e2 = max(e2, -100);            // denormal planes: keep R and 1 / R normal
        const float R = mx > 0.0f ? ldexpf(1.0f, 12 - e2) : 1.0f;
        if (lane == 0) inv[slot] = 1.0f / R;  // a power of two: exact
        const int x = lane & 7, cl = lane >> 3;
        _Float16* hr = Ahi + (2 * cl) * G::AS + G::Cells * slot + x;
        _Float16* lr = Alo + (2 * cl) * G::AS + G::Cells * slot + x;
#pragma unroll
        for (int y = 0; y < EY; y++) {
            const float sc = ((par >> y) & 1) ? -R : R;
            const float re = v[y].x * sc, im = v[y].y * sc;
            const _Float16 rh = (_Float16)re, ih = (_Float16)im;
            hr[8 * y] = rh;
            lr[8 * y] = (_Float16)(re - (float)rh);
            hr[G::AS + 8 * y] = ih;
            lr[G::AS + 8 * y] = (_Float16)(im - (float)ih);
        }
    }
};

// One block of 16 samples' weights (lane = sample l % 16, row group l / 16),
// with the dcw of the sample's entry (dcw_e: per entry, or null).
template <int EY>
struct C2M3Weights {
    using G = C2M3Geom<EY>;
    float4 x0, x1, z0, z1;
    float wy[G::KS], d;
    uint32_t j;
    bool on;

    __device__ __forceinline__ void load(const float* __restrict__ wtab, const uint32_t* __restrict__ list,
                                         const float* __restrict__ dcw_e, uint32_t e0, uint32_t ne, uint32_t bk,
                                         int n, int kg) {
        const uint32_t e = bk * 16 + (uint32_t)n;
        on = e < ne;
        const uint32_t ee = e0 + (on ? e : 0u);  // ne > 0 here
        const float* wr = wtab + (size_t)ee * G::W;
        const float4* w4 = reinterpret_cast<const float4*>(wr);
        x0 = w4[0];
        x1 = w4[1];
        z0 = w4[2 + EY / 4];
        z1 = w4[3 + EY / 4];
#pragma unroll
        for (int h = 0; h < G::KS; h++) wy[h] = wr[8 + 4 * h + kg];
        j = list[ee];
        d = dcw_e ? dcw_e[ee] : 1.0f;
    }
};

// Work-item record of the matrix-core 3D C2NC, one dword per lane (64):
//   [0] first tile's key, [1] number of tiles nt (0: nothing to do),
//   [2 + t] first entry of tile t (t <= nt; [2 + nt] = end), entries of the
//   item's tiles being contiguous in entry order.
constexpr int k3MRec = 64;
constexpr int k3MMaxTiles = k3MRec - 3;

// MODE 0: no dcw, no accumulation; 1: one dcw weight per sample (dcw_e: per
// table entry), no accumulation; 2: anything else.
// Persistent: workgroup b takes items b, b + gridDim.x, ... (gridDim.x a
// multiple of 8 keeps item i on XCD i mod 8, as the items are ordered for).
// While a tile is gathered, the next tile's new planes (one per wave) and the
// record of the item after the current one are in flight; a tile's first
// block of weights per wave is loaded before its planes are split into LDS,
// and the gather itself loads nothing unless a tile has more blocks than
// waves (or MODE 2).  Blocks go to the waves in reverse order, so that a tile
// of fewer new planes than waves keeps every wave busy.  A column's first
// tile has 8 new planes, of which the last four are loaded when the tile is
// staged.
// Three waves per SIMD: registers for the plane and weights in flight (LDS
// would hold four workgroups per CU).
template <int MODE, int EY>
__global__ __launch_bounds__(64 * k3MWaves, C2M3Geom<EY>::Occ) void c2nc_mfma3d_kernel(
    const float* __restrict__ grid, float* __restrict__ out, const float* __restrict__ wtab,
    const uint32_t* __restrict__ list, const uint32_t* __restrict__ recs, int n_items, int M, int F, int C,
    GridGeom g, Layout L, int sign, const float* __restrict__ dcw, long long dcw_len,
    const float* __restrict__ dcw_e, int accumulate, float oscale) {
    using G = C2M3Geom<EY>;
    __shared__ float4 smem[(2 * 16 * G::AS * 2 + k3ME * 4 + 15) / 16];
    _Float16* Ahi = reinterpret_cast<_Float16*>(smem);
    _Float16* Alo = Ahi + 16 * G::AS;
    float* inv = reinterpret_cast<float*>(Alo + 16 * G::AS);  // 1 / R of each plane slot
    const int lane = threadIdx.x & 63, wave = wave_uniform(threadIdx.x >> 6);
    const int c0 = blockIdx.y * k3Coils;
    const int n = lane & 15, kg = lane >> 4;
    const int T = g.tile[2];
    // blocks of 16 samples go to the waves in reverse order (wave 3 first):
    // the new planes of a tile go to the first waves
    const uint32_t wb = (uint32_t)(k3MWaves - 1 - wave);
    // A: row q = l % 16, staged cells (x = 0..7, y = 4 h + kg) of slot s at Cells s + 32 h + 8 kg
    const _Float16* ah = Ahi + n * G::AS + 8 * kg;
    const _Float16* al = Alo + n * G::AS + 8 * kg;
    const long long FM = (long long)F * M;
    float2* o2 = reinterpret_cast<float2*>(out);
    const float* de = MODE == 1 ? dcw_e : nullptr;
    const int stride = (int)gridDim.x;
    auto load_rec = [&](int i) -> uint32_t { return i < n_items ? recs[(size_t)i * k3MRec + lane] : 0u; };
    auto rec_at = [](uint32_t r, int k) -> uint32_t { return (uint32_t)__builtin_amdgcn_readlane((int)r, k); };
    // planes of tile t of a column: t = 0 all 8 from sz0, then T new ones from sz0 + 8 + T (t - 1)
    auto first_plane = [&](const C2M3Column& c, int t) { return t == 0 ? c.sz0 : c.sz0 + k3ME + T * (t - 1); };

    C2M3Plane<EY> pf;  // this wave's first new plane of the step ahead
    // issues the loads of tile t of the column c
    auto prefetch = [&](const C2M3Column& c, int t) {
        const int np = t == 0 ? k3ME : T;
        if (wave < np) pf.load(c, first_plane(c, t) + wave, g, L, sign);
    };

    // the current item (first with tiles), and the record of the next one
    int item = (int)blockIdx.x;
    uint32_t rc = load_rec(item);
    while (item < n_items && rec_at(rc, 1) == 0u) {
        item += stride;
        rc = load_rec(item);
    }
    if (item >= n_items) return;  // uniform over the workgroup; no barrier was passed
    uint32_t rn = load_rec(item + stride);
    C2M3Column col;
    col.setup(grid, rec_at(rc, 0), lane, c0, C, F, g, L);
    int t = 0;
    prefetch(col, 0);
    while (true) {
        const int np = t == 0 ? k3ME : T;
        const int zb = first_plane(col, t);
        const uint32_t e0 = rec_at(rc, 2 + t), ne = rec_at(rc, 3 + t) - e0;
        C2M3Weights<EY> w0;
        if (wb * 16 < ne) w0.load(wtab, list, de, e0, ne, wb, n, kg);
        // split the new planes into their slots (every wave is past the previous gather)
        if (wave < np) pf.store(Ahi, Alo, inv, (zb + wave) & 7, lane);
        for (int j = wave + k3MWaves; j < np; j += k3MWaves) {
            C2M3Plane<EY> q;
            q.load(col, zb + j, g, L, sign);
            q.store(Ahi, Alo, inv, (zb + j) & 7, lane);
        }
        __syncthreads();
        // the step after this one: the next tile of this item, or the next item's first
        int nitem = item, nt = t + 1;
        uint32_t nrc = rc, nrn = rn;
        C2M3Column ncol = col;
        if (nt >= (int)rec_at(rc, 1)) {
            nitem = item + stride;
            nt = 0;
            nrc = rn;
            while (nitem < n_items && rec_at(nrc, 1) == 0u) {  // no-op items pad the ends of the XCD runs
                nitem += stride;
                nrc = load_rec(nitem);
            }
            if (nitem < n_items) {
                nrn = load_rec(nitem + stride);
                ncol.setup(grid, rec_at(nrc, 0), lane, c0, C, F, g, L);
            }
        }
        const bool more = nitem < n_items;
        if (more) prefetch(ncol, nt);

        // gather: staged plane z of this tile is slot (tz0 + z) & 7
        const int tz0 = col.sz0 + T * t;
        const uint32_t nb = (ne + 15) / 16;  // blocks of 16 samples
        for (uint32_t bk = wb; bk < nb; bk += k3MWaves) {
            C2M3Weights<EY> w = w0;
            if (bk != wb) w.load(wtab, list, de, e0, ne, bk, n, kg);  // the first was loaded before
            const mf_float8 wx =
                mf_float8{w.x0.x, w.x0.y, w.x0.z, w.x0.w, w.x1.x, w.x1.y, w.x1.z, w.x1.w} * (w.on ? 1.0f : 0.0f);
            const float wz[k3ME] = {w.z0.x, w.z0.y, w.z0.z, w.z0.w, w.z1.x, w.z1.y, w.z1.z, w.z1.w};
            mf_half8 bh[G::KS], bl[G::KS];
#pragma unroll
            for (int h = 0; h < G::KS; h++) mf_split(wx * w.wy[h], bh[h], bl[h]);
            mf_float4 acc = mf_float4{0.0f, 0.0f, 0.0f, 0.0f};
            // an offset of 0 the compiler cannot see through, renewed per plane:
            // the A rows are read from LDS per block and plane rather than
            // held in 128 registers
            int ao = 0;
#pragma unroll
            for (int z = 0; z < k3ME; z++) {
                asm volatile("" : "+v"(ao));
                const int sl = (tz0 + z) & 7;
                const int o = ao + G::Cells * sl;
                mf_float4 pz = mf_float4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
                for (int h = 0; h < G::KS; h++) {
                    const mf_half8 a_h = *reinterpret_cast<const mf_half8*>(ah + o + 32 * h);
                    const mf_half8 a_l = *reinterpret_cast<const mf_half8*>(al + o + 32 * h);
                    pz = __builtin_amdgcn_mfma_f32_16x16x32_f16(a_l, bh[h], pz, 0, 0, 0);
                    pz = __builtin_amdgcn_mfma_f32_16x16x32_f16(a_h, bl[h], pz, 0, 0, 0);
                    pz = __builtin_amdgcn_mfma_f32_16x16x32_f16(a_h, bh[h], pz, 0, 0, 0);
                }
                acc += (wz[z] * inv[sl]) * pz;  // column l % 16 = this lane's sample
            }
            if (!w.on) continue;
            // accumulator register r: row 4 kg + r = coil 2 kg + r / 2, re / im r % 2
            float d = oscale * w.d;
            const int dm = MODE == 2 ? dcw_mode(dcw, dcw_len, FM) : 0;
            if (MODE == 2 && dm == 1) d *= dcw[w.j];
#pragma unroll
            for (int h = 0; h < 2; h++) {
                const int cc = c0 + 2 * kg + h;
                if (cc >= C) continue;
                float dd = d;
                if (MODE == 2 && dm == 2) dd *= dcw[dcw_index(cc, w.j, FM, dcw_len)];
                float2 val = make_float2(acc[2 * h] * dd, acc[2 * h + 1] * dd);
                float2* o = o2 + (long long)cc * FM + w.j;
                if (MODE == 2 && accumulate) {
                    const float2 old = *o;
                    val.x += old.x;
                    val.y += old.y;
                }
                *o = val;
            }
        }
        if (!more) break;  // uniform; the LDS is not written again
        __syncthreads();   // the gather is done with the planes the next tile overwrites
        item = nitem;
        t = nt;
        rc = nrc;
        rn = nrn;
        col = ncol;
    }
}

// Home tile key of every sample (C2NC binning), any D.
__global__ void home_keys_nd_kernel(const float* __restrict__ p, long long nsamp, long long M, GridGeom g,
                                    uint32_t* __restrict__ keys, uint32_t* __restrict__ vals) {
    const long long j = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nsamp) return;
    uint32_t t = 0, mul = 1;
    for (int d = 0; d < g.D; d++) {
        float pw;
        t += mul * (uint32_t)home_tile_1d(p[j * g.D + d], g, d, pw);
        mul *= (uint32_t)g.ntiles[d];
    }
    keys[j] = (uint32_t)(j / M) * (uint32_t)g.ntiles_total + t;
    vals[j] = (uint32_t)j;
}

}  // namespace gt
