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
//   lanes.
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
// E0 x E1 x E2 (tile + KW - 1 per dimension) as (re, im) pairs in the grid's
// own order (x fastest), then zeroed pad cells -- a window reads up to one cell,
// row and plane past the staged tile (and, with the x halves, 2 XH - KW more
// cells), always with weight 0.  The plane stride is 4 (mod 32) floats: the 8
// coil lanes of one x half, reading (re, im), then fall on banks {4c, 4c + 1},
// and those of the other half, 2 XH floats on (XH = 3 for W 5.5), on the
// remaining 16 banks.
__host__ __device__ constexpr int c2nc3_cells(int KW) { return (k3S0 + KW - 1) * (k3S1 + KW - 1) * (k3S2 + KW - 1); }
__host__ __device__ constexpr int c2nc3_pad_cells(int KW) {
    return (k3S0 + KW - 1) * (k3S1 + KW - 1) + (k3S0 + KW - 1) + 8;
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
    static constexpr int E0 = k3S0 + KW - 1, E1 = k3S1 + KW - 1, E2 = k3S2 + KW - 1;
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
