// gt_kernels_reg.h -- the 2D gridding kernels of the hot path (8 x 8 tiles).
//
// nc2c_reg2d_kernel (NFFT_H_convolve_kernel, ConvolverNC2C_standard.cuh:298-355)
//   One wavefront owns one 8 x 8 tile of the grid for 32 coils; each lane
//   (one coil, real or imaginary part) holds the whole tile in 64 registers.
//   A sample's window sits at one of 15 x 15 positions relative to the tile;
//   a jump on that (wave-uniform) position selects a block of register FMAs
//   specialised at compile time to the in-tile cells of that position, so a
//   sample costs one FMA per in-tile cell and no LDS traffic.  Each lane sums
//   its cells in table order: the result is bitwise repeatable.  LDS is used
//   once per work item, to transpose the tile for coalesced stores.
//
// c2nc_multi2d_kernel (NFFT_convolve_kernel, ConvolverC2NC.cuh:260-307)
//   A workgroup of kWaves wavefronts stages one tile plus the kernel halo
//   (32 coils, centring sign applied) in LDS with batched loads; the
//   wavefronts then take the tile's home samples in turn and gather each
//   window, a whole row of reads before the FMAs.
#pragma once

#include "gt_kernels_fast.h"

namespace gt {

constexpr int kRT = 8;      // register tile edge
constexpr int kPosSpan = 15;  // window-origin positions per dimension: -7 .. 7
constexpr int kWaves = 4;   // wavefronts per C2NC workgroup

template <int KW, int BX, int BY>
__device__ __forceinline__ void reg_update(float (&acc)[kRT * kRT], float s, const float (&wx)[kTabKW],
                                           const float (&wy)[kTabKW]) {
#pragma unroll
    for (int iy = 0; iy < KW; iy++) {
        const int ly = BY + iy;
        if (ly < 0 || ly >= kRT) continue;
        const float sy = s * wy[iy];
#pragma unroll
        for (int ix = 0; ix < KW; ix++) {
            const int lx = BX + ix;
            if (lx < 0 || lx >= kRT) continue;
            acc[ly * kRT + lx] += sy * wx[ix];
        }
    }
}

#define GT_RC(i)                                                                  \
    case (i):                                                                     \
        reg_update<KW, (i) % kPosSpan - 7, (i) / kPosSpan - 7>(acc, s, wx, wy); \
        break;
#define GT_RC5(i) GT_RC(i) GT_RC((i) + 1) GT_RC((i) + 2) GT_RC((i) + 3) GT_RC((i) + 4)
#define GT_RC25(i) GT_RC5(i) GT_RC5((i) + 5) GT_RC5((i) + 10) GT_RC5((i) + 15) GT_RC5((i) + 20)

template <int KW>
__device__ __forceinline__ void reg_update_at(int pos, float (&acc)[kRT * kRT], float s,
                                              const float (&wx)[kTabKW], const float (&wy)[kTabKW]) {
    switch (pos) {
        GT_RC25(0) GT_RC25(25) GT_RC25(50) GT_RC25(75) GT_RC25(100)
        GT_RC25(125) GT_RC25(150) GT_RC25(175) GT_RC25(200)
        default: break;
    }
}
#undef GT_RC25
#undef GT_RC5
#undef GT_RC

template <int KW>
__global__ __launch_bounds__(64) void nc2c_reg2d_kernel(
    const float* __restrict__ samples, float* __restrict__ grid, float* __restrict__ partials,
    const uint32_t* __restrict__ tab, const WorkItem* __restrict__ items, int M, int F, int C, GridGeom g,
    Layout L, int sign, const float* __restrict__ dcw, long long dcw_len, int accumulate) {
    __shared__ float xp[kRT * kRT / 2 * kSubP];  // half-tile transpose buffer
    const int lane = threadIdx.x;
    const WorkItem it = items[blockIdx.x];
    const int c0 = blockIdx.y * kCoilsPerWave;
    const int coil = c0 + (lane >> 1);
    const bool on = coil < C;
    const long long FM = (long long)F * M;
    const float* sl = samples + 2 * (long long)(on ? coil : 0) * FM + (lane & 1);

    float acc[kRT * kRT];
#pragma unroll
    for (int i = 0; i < kRT * kRT; i++) acc[i] = 0.0f;

    for (uint32_t e0 = it.begin; e0 < it.end; e0 += kChunk) {
        const uint32_t ne = min((uint32_t)kChunk, it.end - e0);
        float sv[kChunk];
#pragma unroll
        for (int k = 0; k < kChunk; k++) {
            const uint32_t j = tab[(size_t)(e0 + min((uint32_t)k, ne - 1)) * kTabDw];
            sv[k] = on ? sl[2 * (long long)j] : 0.0f;
            if (dcw && on) sv[k] *= dcw[dcw_index(coil, j, FM, dcw_len)];
        }
#pragma unroll
        for (int k = 0; k < kChunk; k++) {
            if ((uint32_t)k >= ne) break;
            float wx[kTabKW], wy[kTabKW];
            const Entry en = read_entry(tab + (size_t)(e0 + k) * kTabDw, wx, wy);
            const int pos = (en.by + 7) * kPosSpan + (en.bx + 7);
            reg_update_at<KW>(pos, acc, sv[k], wx, wy);
        }
    }

    // write-out through LDS, half a tile (4 rows) at a time: lanes run along x
    // within (coil, y) rows for coalesced stores
    const int f = (int)(it.key / (uint32_t)g.ntiles_total);
    const int t = (int)(it.key % (uint32_t)g.ntiles_total);
    const int ox = (t % g.ntiles[0]) * kRT, oy = (t / g.ntiles[0]) * kRT;
    const int lr = lane >> 4, li = lane & 15;  // 4 rows of 8 cells x (re, im) per pass
    const int lx = li >> 1, comp = li & 1;
    const int gx = ox + lx;
#pragma unroll
    for (int half = 0; half < 2; half++) {
        __syncthreads();
#pragma unroll
        for (int c = 0; c < kRT * kRT / 2; c++) xp[c * kSubP + lane] = acc[half * (kRT * kRT / 2) + c];
        __syncthreads();
        for (int r = lr; r < kCoilsPerWave * (kRT / 2); r += 4) {
            const int cl = r >> 2, ly = half * (kRT / 2) + (r & 3);
            const int cc = c0 + cl;
            const int gy = oy + ly;
            if (cc >= C || gx >= g.nos[0] || gy >= g.nos[1]) continue;
            float v = xp[(((r & 3) * kRT) + lx) * kSubP + 2 * cl + comp];
            if (it.partial >= 0) {
                partials[(((long long)it.partial * C + cc) * (kRT * kRT) + ly * kRT + lx) * 2 + comp] = v;
                continue;
            }
            if (sign && ((gx + gy) & 1)) v = -v;
            const long long o = 2 * ((long long)gx + L.sy * gy + L.sb * ((long long)cc * F + f)) + comp;
            if (accumulate) v += grid[o];
            grid[o] = v;
        }
    }
}

template <int KW>
__global__ __launch_bounds__(64 * kWaves) void c2nc_multi2d_kernel(
    const float* __restrict__ grid, float* __restrict__ out, const uint32_t* __restrict__ tab,
    const WorkItem* __restrict__ items, int M, int F, int C, GridGeom g, Layout L, int sign,
    const float* __restrict__ dcw, long long dcw_len, int accumulate) {
    extern __shared__ float sg[];  // [E1][E0][kSubP] + pad
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const WorkItem it = items[blockIdx.x];
    const int c0 = blockIdx.y * kCoilsPerWave;
    constexpr int E0 = kRT + KW - 1, E1 = kRT + KW - 1;
    const int fr = (int)floorf(g.radius);
    const int f = (int)(it.key / (uint32_t)g.ntiles_total);
    const int t = (int)(it.key % (uint32_t)g.ntiles_total);
    const int sx = (t % g.ntiles[0]) * kRT - fr, sy0 = (t / g.ntiles[0]) * kRT - fr;

    // stage tile + halo: row r = cl * E1 + ey of E0 complex cells; threads run
    // along x, kU rows are loaded before they are stored
    {
        constexpr int rowf = 2 * E0;
        constexpr int rp = (64 * kWaves) / rowf;  // rows per pass
        constexpr int nrows = kCoilsPerWave * E1;
        constexpr int kU = 4;
        const int rs = tid / rowf, q = tid % rowf;
        const int ex = q >> 1, comp = q & 1;
        const int gx = wrap_cell(sx + ex, g.nos[0]);
        if (rs < rp) {
            for (int r0 = rs; r0 < nrows; r0 += rp * kU) {
                float v[kU];
                int dst[kU];
#pragma unroll
                for (int u = 0; u < kU; u++) {
                    const int r = r0 + u * rp;
                    v[u] = 0.0f;
                    dst[u] = -1;
                    if (r < nrows) {
                        const int cl = r / E1, ey = r - (r / E1) * E1;
                        const int gy = wrap_cell(sy0 + ey, g.nos[1]);
                        const int cc = c0 + cl;
                        if (cc < C) {
                            v[u] = grid[2 * ((long long)gx + L.sy * gy + L.sb * ((long long)cc * F + f)) + comp];
                            if (sign && ((gx + gy) & 1)) v[u] = -v[u];
                        }
                        dst[u] = (ey * E0 + ex) * kSubP + 2 * cl + comp;
                    }
                }
#pragma unroll
                for (int u = 0; u < kU; u++)
                    if (dst[u] >= 0) sg[dst[u]] = v[u];
            }
        }
        for (int i = tid; i < KW * kSubP; i += 64 * kWaves) sg[E0 * E1 * kSubP + i] = 0.0f;
    }
    __syncthreads();

    const int coil = c0 + (lane >> 1);
    const bool on = coil < C;
    const long long FM = (long long)F * M;
    float* ol = out + 2 * (long long)(on ? coil : 0) * FM + (lane & 1);
    const float* sl = sg + lane;
    for (uint32_t e0 = it.begin + (uint32_t)wave * kChunk; e0 < it.end; e0 += kChunk * kWaves) {
        const uint32_t ne = min((uint32_t)kChunk, it.end - e0);
        float pre[kChunk];
#pragma unroll
        for (int k = 0; k < kChunk; k++) {
            const uint32_t j = tab[(size_t)(e0 + min((uint32_t)k, ne - 1)) * kTabDw];
            float v = 1.0f;
            if (dcw && on) v = dcw[dcw_index(coil, j, FM, dcw_len)];
            pre[k] = v;
        }
#pragma unroll
        for (int k = 0; k < kChunk; k++) {
            if ((uint32_t)k >= ne) break;
            float wx[kTabKW], wy[kTabKW];
            const Entry en = read_entry(tab + (size_t)(e0 + k) * kTabDw, wx, wy);
            const float* base = sl + (en.by * E0 + en.bx) * kSubP;
            float a = 0.0f;
#pragma unroll
            for (int iy = 0; iy < KW; iy++) {
                if (!((en.my >> iy) & 1u)) continue;
                const float* row = base + iy * E0 * kSubP;
                float v[KW];
#pragma unroll
                for (int ix = 0; ix < KW; ix++) v[ix] = row[ix * kSubP];
                float r = 0.0f;
#pragma unroll
                for (int ix = 0; ix < KW; ix++) r += wx[ix] * v[ix];
                a += wy[iy] * r;
            }
            if (on) {
                const long long o = 2 * (long long)en.j;
                a *= pre[k];
                if (accumulate) a += ol[o];
                ol[o] = a;
            }
        }
    }
}

}  // namespace gt
