// gt_kernels_fast.h -- per-trajectory tables, split-tile reduction and row
// kernels of the 2D hot path (the gridding kernels are in gt_kernels_reg.h).
//
// The gridding kernels run one sample at a time per wavefront, with the 64
// lanes as 32 coils x (real, imaginary), so a sample's control data -- sample
// index, window position, the 2 x KW Kaiser-Bessel weights -- is uniform
// across the wave.  preprocess writes it once per trajectory as a table entry
// per (tile, sample) pair; the kernels read entries with uniform loads, and
// the window bookkeeping runs on the scalar unit.
//
// Grid arrays are addressed through Layout strides, so the same kernels serve
// the internal row-interleaved layout [y][batch][x] of compute() and the
// caller's [x, y, batch] layout of convolve().
#pragma once

#include "gt_kernels.h"

namespace gt {

struct Layout {
    long long sb, sy, sz;  // element strides of batch, y, z; x stride is 1
};

constexpr int kLanes = 64;      // one wavefront
constexpr int kCoilsPerWave = 32;
constexpr int kSubP = 65;       // LDS floats per cell: 64 lanes + 1 (odd stride, conflict-free rows)
constexpr int kTabDw = 16;      // dwords per 2D table entry
constexpr int kTabKW = 7;       // weight slots per dimension in a table entry

// C2NC table entry (16 dwords):
//   [0] sample index j (frame * M + m)
//   [1] packed: bits 0-7 bx (int8), 8-15 by (int8), 16-23 x slot mask, 24-31 y slot mask
//       bx/by = coordinate of window slot 0 in the staged tile (tile origin minus
//       floor(radius)); bit i of a mask marks slot i as inside the window.  Weights of
//       invalid slots are 0.
//   [2..8] x weights of slots 0..6, [9..15] y weights of slots 0..6

__device__ __forceinline__ uint32_t pack_entry(int bx, int by, uint32_t mx, uint32_t my) {
    return (uint32_t)(bx & 0xff) | ((uint32_t)(by & 0xff) << 8) | ((mx & 0xff) << 16) | ((my & 0xff) << 24);
}

// Staged cells per dimension of a C2NC tile of T cells plus the kernel halo.
// Staged offsets start at tile origin - floor(r); a home sample's window
// starts at offset b = 0..T, and its cells of nonzero weight end at offset
// b + 2 floor(r) when r is not an integer (b - 1 + 2r when it is).  With
// KW = floor(2r) + 1 slots, 2 floor(r) is KW - 1 for odd KW (r in [k, k + 1/2))
// and KW - 2 for even KW, so the last cell of nonzero weight is at offset
// T + KW - 1 (odd KW) or T + KW - 2 (even KW).
__host__ __device__ constexpr int c2nc_extent(int T, int KW) { return T + KW - 1 + (KW & 1); }

// Window of one sample along one dimension with unwrapped cell indices:
// slots i = 0..KW-1 are cells lo + i, valid (weight != 0) up to floor(p + r).
template <int KW>
__device__ __forceinline__ void window_raw(float p, const GridGeom& g, int d, float (&w)[KW], int& lo) {
    const float r = g.radius;
    lo = (int)ceilf(p - r);
    const int hi = (int)floorf(p + r);
#pragma unroll
    for (int i = 0; i < KW; i++) {
        const int c = lo + i;
        const float u = fabsf((float)c - p);
        const bool ok = (c <= hi) && !(u > r);
        w[i] = ok ? kb1(u, g.nos_f[d], g.one_over_W, g.beta[d]) : 0.0f;
    }
}

__device__ __forceinline__ int home_tile_1d(float p, const GridGeom& g, int d, float& pw) {
    pw = (p >= g.nos_f[d]) ? p - g.nos_f[d] : p;  // k = +1/2 sits on the cell of k = -1/2
    int c = (int)floorf(pw);
    c = min(max(c, 0), g.nos[d] - 1);
    return c / g.tile[d];
}

// Home tile key (one pair per sample) for the C2NC binning.
__global__ void home_keys_kernel(const float* __restrict__ p, long long nsamp, long long M, GridGeom g,
                                 uint32_t* __restrict__ keys, uint32_t* __restrict__ vals) {
    const long long j = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nsamp) return;
    float pw;
    const int tx = home_tile_1d(p[j * 2 + 0], g, 0, pw);
    const int ty = home_tile_1d(p[j * 2 + 1], g, 1, pw);
    keys[j] = (uint32_t)(j / M) * (uint32_t)g.ntiles_total + (uint32_t)(tx + g.ntiles[0] * ty);
    vals[j] = (uint32_t)j;
}

template <int KW>
__device__ __forceinline__ void write_entry(uint32_t* t, uint32_t j, uint32_t packed, const float (&wx)[KW],
                                            const float (&wy)[KW], uint32_t mx, uint32_t my) {
    t[0] = j;
    t[1] = packed;
#pragma unroll
    for (int i = 0; i < kTabKW; i++) {
        float a = 0.0f, b = 0.0f;
        if (i < KW) {
            a = ((mx >> i) & 1u) ? wx[i < KW ? i : 0] : 0.0f;
            b = ((my >> i) & 1u) ? wy[i < KW ? i : 0] : 0.0f;
        }
        t[2 + i] = __float_as_uint(a);
        t[2 + kTabKW + i] = __float_as_uint(b);
    }
}

// C2NC table: window relative to the staged tile (tile origin - floor(r)).
template <int KW>
__global__ void c2nc_table2d_kernel(const float* __restrict__ p, const uint32_t* __restrict__ list,
                                    long long n, GridGeom g, uint32_t* __restrict__ tab) {
    const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const uint32_t j = list[e];
    const int fr = (int)floorf(g.radius);
    float wx[KW], wy[KW];
    int lox, loy;
    float px, py;
    const int tx = home_tile_1d(p[(long long)j * 2 + 0], g, 0, px);
    const int ty = home_tile_1d(p[(long long)j * 2 + 1], g, 1, py);
    window_raw<KW>(px, g, 0, wx, lox);
    window_raw<KW>(py, g, 1, wy, loy);
    uint32_t mx = 0, my = 0;
#pragma unroll
    for (int i = 0; i < KW; i++) {
        mx |= (wx[i] != 0.0f ? 1u : 0u) << i;
        my |= (wy[i] != 0.0f ? 1u : 0u) << i;
    }
    // 0 <= bx, by <= tile for every home sample (the plan only takes radii away
    // from float rounding of an integer); the clamp keeps reads inside LDS regardless
    const int bx = min(max(lox - (tx * g.tile[0] - fr), 0), g.tile[0]);
    const int by = min(max(loy - (ty * g.tile[1] - fr), 0), g.tile[1]);
    write_entry<KW>(tab + e * kTabDw, j, pack_entry(bx, by, mx, my), wx, wy, mx, my);
}

// Weight rows of the matrix-core C2NC, 32 floats per entry: the sample's x
// weights at the 16 staged columns, then its y weights at the 16 staged rows
// (staged coordinates start at tile origin - floor(r); zero outside the
// window), each axis scaled by a power of two (sx, sy) so no weight exceeds 1.
template <int KW>
__global__ void c2nc_wtable2d_kernel(const float* __restrict__ p, const uint32_t* __restrict__ list, long long n,
                                     GridGeom g, float sx, float sy, float* __restrict__ wtab) {
    const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    static_assert(c2nc_extent(8, KW) <= 16, "a staged row holds 16 cells");
    const uint32_t j = list[e];
    const int fr = (int)floorf(g.radius);
    float wx[KW], wy[KW];
    int lox, loy;
    float px, py;
    const int tx = home_tile_1d(p[(long long)j * 2 + 0], g, 0, px);
    const int ty = home_tile_1d(p[(long long)j * 2 + 1], g, 1, py);
    window_raw<KW>(px, g, 0, wx, lox);
    window_raw<KW>(py, g, 1, wy, loy);
    const int bx = min(max(lox - (tx * g.tile[0] - fr), 0), g.tile[0]);
    const int by = min(max(loy - (ty * g.tile[1] - fr), 0), g.tile[1]);
    float rx[16], ry[16];
#pragma unroll
    for (int c = 0; c < 16; c++) {
        rx[c] = 0.0f;
        ry[c] = 0.0f;
    }
#pragma unroll
    for (int i = 0; i < KW; i++) {
#pragma unroll
        for (int c = 0; c < 16; c++) {
            if (c == bx + i) rx[c] = wx[i] * sx;
            if (c == by + i) ry[c] = wy[i] * sy;
        }
    }
    float4* t = reinterpret_cast<float4*>(wtab + e * 32);
#pragma unroll
    for (int q = 0; q < 4; q++) {
        t[q] = make_float4(rx[4 * q], rx[4 * q + 1], rx[4 * q + 2], rx[4 * q + 3]);
        t[4 + q] = make_float4(ry[4 * q], ry[4 * q + 1], ry[4 * q + 2], ry[4 * q + 3]);
    }
}

// NC2C table, per (8 x 8 tile, sample) pair e, in two arrays:
//   wt[16 e + c]     weight of tile column c (c = 0..7), then of tile row c - 8
//                    (zero where the window does not reach)
//   hd[2 e]          sample index j
//   hd[2 e + 1]      bits 0-7: tile rows the window covers, bits 8-15: tile columns
// Cells are matched by their wrapped coordinate, so windows that wrap around the
// grid edge need no special case.
constexpr int kTabNW = 16;  // weight floats per NC2C entry (one 64-byte line)
constexpr int kTabTile = 8;

// The column weights are stored times csc, a power of two (the matrix-core
// kernel's weight normaliser, so that it reads them ready to multiply; the
// FP32 kernel scales its sums back by 1 / csc, exactly).
template <int KW>
__global__ void nc2c_table2d_kernel(const float* __restrict__ p, const uint32_t* __restrict__ list,
                                    const uint32_t* __restrict__ keys, long long n, GridGeom g,
                                    float* __restrict__ wt, uint32_t* __restrict__ hd, float csc) {
    const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const uint32_t j = list[e];
    const int tile = (int)(keys[e] % (uint32_t)g.ntiles_total);
    const int org[2] = {(tile % g.ntiles[0]) * g.tile[0], (tile / g.ntiles[0]) * g.tile[1]};
    float tw[2][kTabTile];
    uint32_t mask[2] = {0u, 0u};  // tile columns / rows the window covers
    for (int d = 0; d < 2; d++) {
        float w[KW];
        int lo;
        window_raw<KW>(p[(long long)j * 2 + d], g, d, w, lo);
#pragma unroll
        for (int c = 0; c < kTabTile; c++) tw[d][c] = 0.0f;
        for (int i = 0; i < KW; i++) {
            if (w[i] == 0.0f) continue;
            const int l = wrap_cell(lo + i, g.nos[d]) - org[d];
            if (l >= 0 && l < kTabTile) {
                tw[d][l] = d == 0 ? w[i] * csc : w[i];
                mask[d] |= 1u << l;
            }
        }
    }
    const uint32_t rowmask = mask[1] | (mask[0] << 8);
    float4* t = reinterpret_cast<float4*>(wt + e * kTabNW);
    t[0] = make_float4(tw[0][0], tw[0][1], tw[0][2], tw[0][3]);
    t[1] = make_float4(tw[0][4], tw[0][5], tw[0][6], tw[0][7]);
    t[2] = make_float4(tw[1][0], tw[1][1], tw[1][2], tw[1][3]);
    t[3] = make_float4(tw[1][4], tw[1][5], tw[1][6], tw[1][7]);
    hd[2 * e] = j;
    hd[2 * e + 1] = rowmask;
}

struct Entry {
    uint32_t j;
    int bx, by;
    uint32_t mx, my;
};

__device__ __forceinline__ Entry read_entry(const uint32_t* __restrict__ t, float* wx, float* wy) {
    Entry en;
    en.j = t[0];
    const uint32_t pk = t[1];
    en.bx = (int)(int8_t)(pk & 0xff);
    en.by = (int)(int8_t)((pk >> 8) & 0xff);
    en.mx = (pk >> 16) & 0xff;
    en.my = pk >> 24;
#pragma unroll
    for (int i = 0; i < kTabKW; i++) {
        wx[i] = __uint_as_float(t[2 + i]);
        wy[i] = __uint_as_float(t[2 + kTabKW + i]);
    }
    return en;
}

// Index into dcw for sample element (coil, j): the reference broadcasts dcw by
// modulus of its length (cuNDArray_operators.cu:129-145).
__device__ __forceinline__ long long dcw_index(int coil, uint32_t j, long long FM, long long dcw_len) {
    if (dcw_len == FM) return j;
    const long long i = (long long)coil * FM + j;
    return i < dcw_len ? i : i % dcw_len;
}

// How a kernel applies dcw, uniform over the launch: 0 none, 1 one weight per
// sample shared by every coil (a wave-uniform load), 2 per coil element.
__device__ __forceinline__ int dcw_mode(const float* dcw, long long dcw_len, long long FM) {
    return !dcw ? 0 : (dcw_len == FM ? 1 : 2);
}

// LDS floats of the C2NC staged tile: c2nc_extent(T1) rows of c2nc_extent(T0)
// cells, then one zeroed row and one zeroed cell: reading all KW x KW slots
// from offset b <= T reaches at most one row and one cell past the staged
// tile (even KW), always in slots of weight 0.
// Rounded up to whole float4s: the work item's table entries follow it.
// A staged cell holds the 2 x 32 / G floats of a wave's coils plus one (an
// odd stride keeps rows conflict-free).
// LDS floats per staged cell of the 2D C2NC: 32 / G coils x (re, im) plus a
// pad float; with 32 coils (G = 1) a lane reads a coil's (re, im) as one
// 8-byte word, so the stride stays even.
__host__ __device__ constexpr int c2nc_cell_stride(int G) {
    return G == 1 ? 2 * kCoilsPerWave + 2 : 2 * (kCoilsPerWave / G) + 1;
}

__host__ __device__ inline int c2nc_lds_floats(int T0, int T1, int KW, int G = 1) {
    return (((c2nc_extent(T1, KW) + 1) * c2nc_extent(T0, KW) + 1) * c2nc_cell_stride(G) + 3) & ~3;
}

// C2NC work items hold at most this many entries (their table is staged in LDS).
constexpr int kC2NCSeg = 128;

__host__ __device__ inline int c2nc_lds_bytes(int KW, int G = 1) {
    return (c2nc_lds_floats(8, 8, KW, G) + kC2NCSeg * kTabDw + kC2NCSeg) * 4;
}

// ---------------------------------------------------------------------------
// Row kernels between the caller's image layout [x, y, batch] and the internal
// grid layout [y][batch][x]: one workgroup per grid (or image) row.
// ---------------------------------------------------------------------------
struct RowGeom {
    int Nx, Ny;       // image size
    int Ox, Oy;       // oversampled size
    int offx, offy;   // pad or crop offsets
    int nb;           // batches (F * C)
    const float* fx;  // reciprocal 1-D deapodization filters
    const float* fy;
    float scale;
    int sign;
};

// grid row r = y * nb + b (internal layout); writes the full row.
__global__ __launch_bounds__(256) void pad_rows2d_kernel(const float2* __restrict__ img,
                                                         float2* __restrict__ grid, RowGeom rg) {
    const int r = blockIdx.x;
    const int y = r / rg.nb, b = r % rg.nb;
    const int iy = y - rg.offy;
    float2* out = grid + (long long)r * rg.Ox;
    if (iy < 0 || iy >= rg.Ny) {
        for (int x = threadIdx.x; x < rg.Ox; x += blockDim.x) out[x] = make_float2(0.0f, 0.0f);
        return;
    }
    const float2* in = img + ((long long)b * rg.Ny + iy) * rg.Nx;
    const float sy = rg.scale * rg.fy[y] * ((rg.sign && (y & 1)) ? -1.0f : 1.0f);
    for (int x = threadIdx.x; x < rg.Ox; x += blockDim.x) {
        const int ix = x - rg.offx;
        float2 v = make_float2(0.0f, 0.0f);
        if (ix >= 0 && ix < rg.Nx) {
            v = in[ix];
            const float s = sy * rg.fx[x] * ((rg.sign && (x & 1)) ? -1.0f : 1.0f);
            v.x *= s;
            v.y *= s;
        }
        out[x] = v;
    }
}

// image row r = b * Ny + iy; reads grid row (iy + offy) * nb + b.
__global__ __launch_bounds__(256) void crop_rows2d_kernel(const float2* __restrict__ grid,
                                                          float2* __restrict__ img, RowGeom rg) {
    const int r = blockIdx.x;
    const int b = r / rg.Ny, iy = r % rg.Ny;
    const int y = iy + rg.offy;
    const float2* in = grid + ((long long)y * rg.nb + b) * rg.Ox;
    float2* out = img + (long long)r * rg.Nx;
    const float sy = rg.scale * rg.fy[y] * ((rg.sign && (y & 1)) ? -1.0f : 1.0f);
    for (int ix = threadIdx.x; ix < rg.Nx; ix += blockDim.x) {
        const int x = ix + rg.offx;
        float2 v = in[x];
        const float s = sy * rg.fx[x] * ((rg.sign && (x & 1)) ? -1.0f : 1.0f);
        v.x *= s;
        v.y *= s;
        out[ix] = v;
    }
}

}  // namespace gt
