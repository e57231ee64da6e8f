// gt_fft.h -- the 2D and 3D centred FFTs of the hot path, fused with pad / crop and
// deapodization (cuNFFT_impl::fft + deapodize, cuNFFT.cu:57-92, and the pad /
// crop passes of NFFT_plan::compute, NFFT.hpp:125-184).
//
// The grid is the row-interleaved layout [y][batch][x] of compute().  Of the
// oversampled grid only the band of rows holding the image (forward) or kept
// by the crop (backward) carries information, so the 2D FFT runs as
//   forward : image rows --pad, deapodize, sign, scale, row FFT--> band rows
//             band rows (others are zero) --column FFT--> all rows
//   backward: all rows --column FFT--> band rows
//             band rows --row FFT, crop, deapodize, sign, scale--> image rows
// and the rows outside the band are neither written nor read.  The centring
// sign (-1)^(x+y) is applied by the neighbouring kernels (pad / crop here, the
// gridding kernels on the grid side), so the transforms are plain DFTs,
// e^{-2 pi i k n / L} forwards and e^{+2 pi i k n / L} backwards, unscaled.
//
// Each 1D transform is a Stockham autosort FFT in LDS (radices 8, 4, 2, 3, 5,
// chosen at compile time), in place through registers: a pass reads its
// butterfly inputs, synchronizes, and writes its outputs.  Twiddles come from
// a per-length table exp(-2 pi i m / L), m < L, built on the host in double.
#pragma once

#include <hip/hip_runtime.h>

#include "gt_kernels_fast.h"

namespace gt {

constexpr int fft_next_radix(int m) {
    return (m % 8 == 0) ? 8 : (m % 4 == 0) ? 4 : (m % 2 == 0) ? 2 : (m % 3 == 0) ? 3 : (m % 5 == 0) ? 5 : 0;
}

// True if L factors into the radices above (and is one of the compiled lengths).
constexpr bool fft_factorable(int L) {
    int m = L;
    while (m > 1) {
        const int r = fft_next_radix(m);
        if (r == 0) return false;
        m /= r;
    }
    return true;
}

__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
// multiply by -i (s = -1, forwards) or +i (s = +1, backwards)
__device__ __forceinline__ float2 cmul_i(float2 a, float s) { return make_float2(-s * a.y, s * a.x); }

// In-register DFT of R points; s = -1 forwards, +1 backwards.
template <int R>
__device__ __forceinline__ void dft(float2 (&x)[R], float s);

template <>
__device__ __forceinline__ void dft<2>(float2 (&x)[2], float) {
    const float2 a = x[0], b = x[1];
    x[0] = cadd(a, b);
    x[1] = csub(a, b);
}

template <>
__device__ __forceinline__ void dft<3>(float2 (&x)[3], float s) {
    // w = e^{s 2 pi i / 3} = -1/2 + s i sqrt(3)/2
    const float c = -0.5f, d = s * 0.86602540378443864676f;
    const float2 t = cadd(x[1], x[2]);
    const float2 u = csub(x[1], x[2]);
    const float2 m = make_float2(x[0].x + c * t.x, x[0].y + c * t.y);
    const float2 n = make_float2(-d * u.y, d * u.x);  // i d u
    x[0] = cadd(x[0], t);
    x[1] = cadd(m, n);
    x[2] = csub(m, n);
}

template <>
__device__ __forceinline__ void dft<4>(float2 (&x)[4], float s) {
    const float2 a = cadd(x[0], x[2]), b = csub(x[0], x[2]);
    const float2 c = cadd(x[1], x[3]), d = cmul_i(csub(x[1], x[3]), -s);
    // X1 = b + w d' with w^1 = e^{s i pi/2} = s i:  (x1 - x3) * s i
    x[0] = cadd(a, c);
    x[2] = csub(a, c);
    x[1] = csub(b, d);
    x[3] = cadd(b, d);
}

template <>
__device__ __forceinline__ void dft<5>(float2 (&x)[5], float s) {
    // cos/sin of 2 pi / 5 and 4 pi / 5
    const float c1 = 0.30901699437494742410f, c2 = -0.80901699437494742410f;
    const float s1 = s * 0.95105651629515357212f, s2 = s * 0.58778525229247312917f;
    const float2 t1 = cadd(x[1], x[4]), t2 = cadd(x[2], x[3]);
    const float2 u1 = csub(x[1], x[4]), u2 = csub(x[2], x[3]);
    const float2 a1 = make_float2(x[0].x + c1 * t1.x + c2 * t2.x, x[0].y + c1 * t1.y + c2 * t2.y);
    const float2 a2 = make_float2(x[0].x + c2 * t1.x + c1 * t2.x, x[0].y + c2 * t1.y + c1 * t2.y);
    // i (s1 u1 + s2 u2) and i (s2 u1 - s1 u2)
    const float2 b1 = make_float2(-(s1 * u1.y + s2 * u2.y), s1 * u1.x + s2 * u2.x);
    const float2 b2 = make_float2(-(s2 * u1.y - s1 * u2.y), s2 * u1.x - s1 * u2.x);
    x[0] = cadd(x[0], cadd(t1, t2));
    x[1] = cadd(a1, b1);
    x[4] = csub(a1, b1);
    x[2] = cadd(a2, b2);
    x[3] = csub(a2, b2);
}

template <>
__device__ __forceinline__ void dft<8>(float2 (&x)[8], float s) {
    // radix-2 over even/odd, then radix-4 halves
    float2 e[4] = {x[0], x[2], x[4], x[6]};
    float2 o[4] = {x[1], x[3], x[5], x[7]};
    dft<4>(e, s);
    dft<4>(o, s);
    const float h = 0.70710678118654752440f;
    // o[k] *= e^{s 2 pi i k / 8}
    o[1] = make_float2(h * (o[1].x - s * o[1].y), h * (o[1].y + s * o[1].x));
    o[2] = cmul_i(o[2], s);
    o[3] = make_float2(h * (-o[3].x - s * o[3].y), h * (-o[3].y + s * o[3].x));
#pragma unroll
    for (int k = 0; k < 4; k++) {
        x[k] = cadd(e[k], o[k]);
        x[k + 4] = csub(e[k], o[k]);
    }
}

// LDS slot of element i of a sequence.  Contiguous sequences (row
// transforms, ROW) keep each aligned block of 16 elements (one pass of the 32
// four-byte banks for 8-byte accesses) and permute it: the low four bits are
// XORed with g(i / 16) = (i / 16) mod 8 + 8 ((i / 16) & 4 != 0).  Contiguous
// reads stay one block per 16 lanes; the stride-8 writes of a first radix-8
// pass (16 lanes over 8 blocks) and the split writes of a second (two runs of
// 8, 4 blocks apart) reach 16 distinct banks pairs.  Interleaved sequences
// (columns) use a plain stride.
template <bool ROW>
__device__ __forceinline__ int fft_slot(int i, int estride) {
    if (ROW) {
        const int r = i >> 4;
        return i ^ ((r & 7) | ((r & 4) << 1));
    }
    return i * estride;
}
template <int L, bool ROW>
constexpr int fft_slots() {
    return ROW ? (L + 15) / 16 * 16 : L;
}

// One Stockham pass of radix R on sequences of length L held in LDS (element
// i of this thread's sequence at buf[fft_slot(i)]).  NS = product of the
// radices already applied; t = 0..TPS-1 is the thread's index within its
// sequence.  Every thread of the block takes part (two barriers).
template <int L, int R, int NS, int TPS, bool PADR>
__device__ __forceinline__ void fft_pass(float2* buf, int estride, int t, const float2* __restrict__ tw, float s) {
    constexpr int NB = L / R;
    constexpr int PER = (NB + TPS - 1) / TPS;
    float2 x[PER][R];
#pragma unroll
    for (int i = 0; i < PER; i++) {
        const int j = t + i * TPS;
        if (j < NB) {
#pragma unroll
            for (int q = 0; q < R; q++) x[i][q] = buf[fft_slot<PADR>(j + q * NB, estride)];
        }
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < PER; i++) {
        const int j = t + i * TPS;
        if (j < NB) {
            const int k = j % NS;
            if (NS > 1) {
#pragma unroll
                for (int q = 1; q < R; q++) {
                    // k q < NS R: index below L; the table holds its first half
                    // (e^{-i theta} for m < L / 2; the second half is its negative)
                    const int m = k * q * (L / (NS * R));
                    const bool up = m >= L / 2;
                    float2 w = tw[up ? m - L / 2 : m];
                    if (up) w = make_float2(-w.x, -w.y);
                    w.y *= -s;  // backwards conjugates
                    x[i][q] = cmul(x[i][q], w);
                }
            }
            dft<R>(x[i], s);
            const int base = (j / NS) * NS * R + k;
#pragma unroll
            for (int q = 0; q < R; q++) buf[fft_slot<PADR>(base + q * NS, estride)] = x[i][q];
        }
    }
    __syncthreads();
}

template <int L, int NS, int TPS, bool PADR>
__device__ __forceinline__ void fft_passes(float2* buf, int estride, int t, const float2* __restrict__ tw, float s) {
    if constexpr (NS < L) {
        constexpr int R = fft_next_radix(L / NS);
        static_assert(R > 0, "FFT length must factor into 2, 3, 5");
        fft_pass<L, R, NS, TPS, PADR>(buf, estride, t, tw, s);
        fft_passes<L, NS * R, TPS, PADR>(buf, estride, t, tw, s);
    }
}

// First half of the twiddle table of length L (even) into LDS, all threads of
// the block: e^{-2 pi i (m + L/2) / L} = -e^{-2 pi i m / L}, so the passes
// need only m < L / 2 (half the LDS: the column transforms of length 384 fit
// six workgroups per CU instead of five).
template <int L>
__device__ __forceinline__ void load_twiddles(float2* sw, const float2* __restrict__ tw) {
    static_assert(L % 2 == 0, "even transform length");
    for (int i = threadIdx.x; i < L / 2; i += blockDim.x) sw[i] = tw[i];
}

struct FftGeom {
    int Nx, Ny;         // image size
    int offx, offy;     // band offset (pad offset forwards, crop offset backwards)
    int nb;             // batches F * C
    const float* fx;    // reciprocal deapodization filters (length Ox, Oy)
    const float* fy;
    float scale;        // 1/sqrt(G) times the centring global sign
    // forwards only: coil maps [Ny][Nx] per coil (or null); batch b = c * F + f
    // then reads image f times map c (the SENSE coil multiply)
    const float2* csm;
    int F;
    // forwards, many problems: the image and coil map of batch b are
    // img_b[b] and csm_b[b] (each [Ny][Nx]); null: img / csm above
    const float2* const* img_b;
    const float2* const* csm_b;
};

constexpr int kFftRowThreads = 64;  // threads per row transform
constexpr int kFftRows = 4;         // rows per workgroup
constexpr int kFftColW = 16;        // columns per workgroup (one 128-byte run of a row)

// Columns per workgroup of the column FFT of length OY: 8 (a 64-byte half run,
// half the LDS, so more workgroups per CU) where 32 threads per column divide
// the length, else 16.
template <int OY>
__host__ __device__ constexpr int fft_col_width() { return OY % 32 == 0 ? 8 : kFftColW; }

// Column block of launch index i: with 8-column blocks the two halves of a
// 128-byte run are blocks i and i + 8, which run on the same XCD (workgroup i
// goes to XCD i mod 8) and share its L2.  Returns false past the last block.
template <int W>
__device__ __forceinline__ bool fft_col_block(int i, int nb, int nxb, int& b, int& xb) {
    int lin = i;
    if (W == 8) lin = ((i >> 4) * 8 + (i & 7)) * 2 + ((i >> 3) & 1);
    if (lin >= nb * nxb) return false;
    b = lin / nxb;
    xb = lin % nxb;
    return true;
}

// Forwards: image row (b, iy) -> grid row (iy + offy) * nb + b, padded to Ox
// at offx, deapodized, centring-signed and scaled, then transformed along x.
template <int OX>
__global__ __launch_bounds__(kFftRowThreads* kFftRows) void fft_rows_pad_kernel(
    const float2* __restrict__ img, float2* __restrict__ grid, FftGeom fg, const float2* __restrict__ tw,
    float s) {
    constexpr int NS = fft_slots<OX, true>();
    __shared__ float2 sbuf[kFftRows][NS];
    __shared__ float2 stw[OX / 2];
    load_twiddles<OX>(stw, tw);
    const int rl = threadIdx.x / kFftRowThreads, t = threadIdx.x % kFftRowThreads;
    const int r = blockIdx.x * kFftRows + rl;  // image row b * Ny + iy
    const int nrows = fg.nb * fg.Ny;
    const bool live = r < nrows;
    const int b = live ? r / fg.Ny : 0, iy = live ? r % fg.Ny : 0;
    const int y = iy + fg.offy;
    const float sy = fg.scale * fg.fy[y] * ((y & 1) ? -1.0f : 1.0f);
    const bool tables = fg.img_b != nullptr;  // block-uniform
    const bool coils = tables ? fg.csm_b != nullptr : fg.csm != nullptr;
    const float2* in;
    const float2* cm = nullptr;
    if (tables) {
        in = fg.img_b[b] + (long long)iy * fg.Nx;
        if (coils) cm = fg.csm_b[b] + (long long)iy * fg.Nx;
    } else {
        const int fb = coils ? b % fg.F : b;  // image batch read
        in = img + ((long long)fb * fg.Ny + iy) * fg.Nx;
        if (coils) cm = fg.csm + ((long long)(b / fg.F) * fg.Ny + iy) * fg.Nx;
    }
    float2* sb = &sbuf[rl][0];
    // every load of the row before the first LDS store (one memory latency)
    constexpr int kPer = (OX + kFftRowThreads - 1) / kFftRowThreads;
    float2 v[kPer];
    float fxs[kPer];
#pragma unroll
    for (int i = 0; i < kPer; i++) {
        const int x = t + i * kFftRowThreads;
        const int ix = x - fg.offx;
        const bool ok = live && x < OX && ix >= 0 && ix < fg.Nx;
        v[i] = ok ? in[ix] : make_float2(0.0f, 0.0f);
        fxs[i] = x < OX ? fg.fx[x] : 0.0f;
    }
    if (coils) {
#pragma unroll
        for (int i = 0; i < kPer; i++) {
            const int x = t + i * kFftRowThreads;
            const int ix = x - fg.offx;
            const bool ok = live && x < OX && ix >= 0 && ix < fg.Nx;
            const float2 c = ok ? cm[ix] : make_float2(0.0f, 0.0f);
            v[i] = make_float2(v[i].x * c.x - v[i].y * c.y, v[i].x * c.y + v[i].y * c.x);
        }
    }
#pragma unroll
    for (int i = 0; i < kPer; i++) {
        const int x = t + i * kFftRowThreads;
        if (x < OX) {
            const float f = sy * fxs[i] * ((x & 1) ? -1.0f : 1.0f);
            sb[fft_slot<true>(x, 1)] = make_float2(v[i].x * f, v[i].y * f);
        }
    }
    __syncthreads();
    fft_passes<OX, 1, kFftRowThreads, true>(sb, 1, t, stw, s);
    if (!live) return;
    float2* out = grid + ((long long)y * fg.nb + b) * OX;
    for (int x = t; x < OX; x += kFftRowThreads) out[x] = sb[fft_slot<true>(x, 1)];
}

// Backwards: grid row (iy + offy) * nb + b, transformed along x, cropped at
// offx, deapodized, centring-signed and scaled -> image row (b, iy).
template <int OX>
__global__ __launch_bounds__(kFftRowThreads* kFftRows) void fft_rows_crop_kernel(
    const float2* __restrict__ grid, float2* __restrict__ img, FftGeom fg, const float2* __restrict__ tw,
    float s) {
    constexpr int NS = fft_slots<OX, true>();
    __shared__ float2 sbuf[kFftRows][NS];
    __shared__ float2 stw[OX / 2];
    load_twiddles<OX>(stw, tw);
    const int rl = threadIdx.x / kFftRowThreads, t = threadIdx.x % kFftRowThreads;
    const int r = blockIdx.x * kFftRows + rl;
    const int nrows = fg.nb * fg.Ny;
    const bool live = r < nrows;
    const int b = live ? r / fg.Ny : 0, iy = live ? r % fg.Ny : 0;
    const int y = iy + fg.offy;
    const float2* in = grid + ((long long)y * fg.nb + b) * OX;
    float2* sb = &sbuf[rl][0];
    {
        constexpr int kPer = (OX + kFftRowThreads - 1) / kFftRowThreads;
        float2 v[kPer];
#pragma unroll
        for (int i = 0; i < kPer; i++) {
            const int x = t + i * kFftRowThreads;
            v[i] = (live && x < OX) ? in[x] : make_float2(0.0f, 0.0f);
        }
#pragma unroll
        for (int i = 0; i < kPer; i++) {
            const int x = t + i * kFftRowThreads;
            if (x < OX) sb[fft_slot<true>(x, 1)] = v[i];
        }
    }
    __syncthreads();
    fft_passes<OX, 1, kFftRowThreads, true>(sb, 1, t, stw, s);
    if (!live) return;
    const float sy = fg.scale * fg.fy[y] * ((y & 1) ? -1.0f : 1.0f);
    float2* out = img + ((long long)b * fg.Ny + iy) * fg.Nx;
    for (int ix = t; ix < fg.Nx; ix += kFftRowThreads) {
        const int x = ix + fg.offx;
        float2 v = sb[fft_slot<true>(x, 1)];
        const float f = sy * fg.fx[x] * ((x & 1) ? -1.0f : 1.0f);
        v.x *= f;
        v.y *= f;
        out[ix] = v;
    }
}

// Column transforms along y for kFftColW consecutive columns of one batch.
// BAND_IN: rows outside [offy, offy + Ny) are zero (not read).
// BAND_OUT: only rows inside [offy, offy + Ny) are written.
template <int OY, bool BAND_IN, bool BAND_OUT>
__global__ __launch_bounds__(256) void fft_cols_kernel(float2* __restrict__ grid, int Ox, FftGeom fg,
                                                       const float2* __restrict__ tw, float s) {
    constexpr int W = fft_col_width<OY>();
    constexpr int TPS = 256 / W;
    __shared__ float2 sbuf[OY * W];
    __shared__ float2 stw[OY / 2];
    const int nxb = (Ox + W - 1) / W;
    int b, xb;
    if (!fft_col_block<W>((int)blockIdx.x, fg.nb, nxb, b, xb)) return;  // block-uniform, before any barrier
    load_twiddles<OY>(stw, tw);
    const int x0 = xb * W;
    const int xl = threadIdx.x % W, yr = threadIdx.x / W;
    const int x = x0 + xl;
    const bool xin = x < Ox;
    const long long rs = (long long)fg.nb * Ox;  // grid row stride
    float2* col = grid + (long long)b * Ox + x;
    const int ylo = fg.offy, yhi = fg.offy + fg.Ny;
    {
        // every load of the column block before the first LDS store
        static_assert(OY % TPS == 0, "column length must be a multiple of the threads per column");
        constexpr int kPer = OY / TPS;
        float2 v[kPer];
#pragma unroll
        for (int i = 0; i < kPer; i++) {
            const int y = yr + i * TPS;
            v[i] = (xin && (!BAND_IN || (y >= ylo && y < yhi))) ? col[y * rs] : make_float2(0.0f, 0.0f);
        }
#pragma unroll
        for (int i = 0; i < kPer; i++) sbuf[(yr + i * TPS) * W + xl] = v[i];
    }
    __syncthreads();
    fft_passes<OY, 1, TPS, false>(&sbuf[xl], W, yr, stw, s);
    if (!xin) return;
    for (int y = yr; y < OY; y += TPS) {
        if (BAND_OUT && (y < ylo || y >= yhi)) continue;
        col[y * rs] = sbuf[y * W + xl];
    }
}

// ---- 3D band transforms ---------------------------------------------------
// The 3D grid has x contiguous and strides sb, sy, sz (in cells) for batch, y
// and z (the gridding kernels' Layout).  The image occupies the band
// [off, off + N) of every dimension, so
//   forward : image lines (b, iz, iy) --pad, deapodize, sign, scale, x FFT-->
//             grid lines (b, z, y) of the y and z bands;
//             y FFT of the planes of the z band (y outside its band is zero);
//             z FFT of every (y, x) line (z outside its band is zero)
//   backward: z FFT of every (y, x) line, writing the z band;
//             y FFT of the planes of the z band, writing the y band;
//             x FFT of the lines of both bands --crop, deapodize, sign, scale-->
//             image lines (b, iz, iy)
// Lines and planes outside the bands are neither read nor written.
struct Fft3Geom {
    int Nx, Ny, Nz;        // image size
    int Ox, Oy, Oz;        // grid size
    int offx, offy, offz;  // band offsets
    int nb;                // batches
    long long sb, sy, sz;  // grid strides of batch, y, z
    const float* fx;       // reciprocal deapodization filters (length Ox, Oy, Oz)
    const float* fy;
    const float* fz;
    float scale;           // 1/sqrt(G) times the centring global sign
};

// Forwards: image line r = (b * Nz + iz) * Ny + iy -> grid line (b, iz + offz,
// iy + offy), padded at offx, deapodized, centring-signed, scaled, transformed.
template <int OX>
__global__ __launch_bounds__(kFftRowThreads* kFftRows) void fft3_rows_pad_kernel(
    const float2* __restrict__ img, float2* __restrict__ grid, Fft3Geom fg, const float2* __restrict__ tw,
    float s) {
    constexpr int NS = fft_slots<OX, true>();
    __shared__ float2 sbuf[kFftRows][NS];
    __shared__ float2 stw[OX / 2];
    load_twiddles<OX>(stw, tw);
    const int rl = threadIdx.x / kFftRowThreads, t = threadIdx.x % kFftRowThreads;
    const long long r = (long long)blockIdx.x * kFftRows + rl;
    const bool live = r < (long long)fg.nb * fg.Nz * fg.Ny;
    const long long rr = live ? r : 0;
    const int iy = (int)(rr % fg.Ny), iz = (int)((rr / fg.Ny) % fg.Nz);
    const long long b = rr / ((long long)fg.Ny * fg.Nz);
    const int y = iy + fg.offy, z = iz + fg.offz;
    const float syz = fg.scale * fg.fy[y] * fg.fz[z] * (((y + z) & 1) ? -1.0f : 1.0f);
    const float2* in = img + rr * fg.Nx;
    float2* sb = &sbuf[rl][0];
    constexpr int kPer = (OX + kFftRowThreads - 1) / kFftRowThreads;
    float2 v[kPer];
    float fxs[kPer];
#pragma unroll
    for (int i = 0; i < kPer; i++) {
        const int x = t + i * kFftRowThreads;
        const int ix = x - fg.offx;
        const bool ok = live && x < OX && ix >= 0 && ix < fg.Nx;
        v[i] = ok ? in[ix] : make_float2(0.0f, 0.0f);
        fxs[i] = x < OX ? fg.fx[x] : 0.0f;
    }
#pragma unroll
    for (int i = 0; i < kPer; i++) {
        const int x = t + i * kFftRowThreads;
        if (x < OX) {
            const float f = syz * fxs[i] * ((x & 1) ? -1.0f : 1.0f);
            sb[fft_slot<true>(x, 1)] = make_float2(v[i].x * f, v[i].y * f);
        }
    }
    __syncthreads();
    fft_passes<OX, 1, kFftRowThreads, true>(sb, 1, t, stw, s);
    if (!live) return;
    float2* out = grid + b * fg.sb + y * fg.sy + z * fg.sz;
    for (int x = t; x < OX; x += kFftRowThreads) out[x] = sb[fft_slot<true>(x, 1)];
}

// Backwards: grid line (b, iz + offz, iy + offy), transformed along x, cropped
// at offx, deapodized, centring-signed, scaled -> image line r.
template <int OX>
__global__ __launch_bounds__(kFftRowThreads* kFftRows) void fft3_rows_crop_kernel(
    const float2* __restrict__ grid, float2* __restrict__ img, Fft3Geom fg, const float2* __restrict__ tw,
    float s) {
    constexpr int NS = fft_slots<OX, true>();
    __shared__ float2 sbuf[kFftRows][NS];
    __shared__ float2 stw[OX / 2];
    load_twiddles<OX>(stw, tw);
    const int rl = threadIdx.x / kFftRowThreads, t = threadIdx.x % kFftRowThreads;
    const long long r = (long long)blockIdx.x * kFftRows + rl;
    const bool live = r < (long long)fg.nb * fg.Nz * fg.Ny;
    const long long rr = live ? r : 0;
    const int iy = (int)(rr % fg.Ny), iz = (int)((rr / fg.Ny) % fg.Nz);
    const long long b = rr / ((long long)fg.Ny * fg.Nz);
    const int y = iy + fg.offy, z = iz + fg.offz;
    const float2* in = grid + b * fg.sb + y * fg.sy + z * fg.sz;
    float2* sb = &sbuf[rl][0];
    {
        constexpr int kPer = (OX + kFftRowThreads - 1) / kFftRowThreads;
        float2 v[kPer];
#pragma unroll
        for (int i = 0; i < kPer; i++) {
            const int x = t + i * kFftRowThreads;
            v[i] = (live && x < OX) ? in[x] : make_float2(0.0f, 0.0f);
        }
#pragma unroll
        for (int i = 0; i < kPer; i++) {
            const int x = t + i * kFftRowThreads;
            if (x < OX) sb[fft_slot<true>(x, 1)] = v[i];
        }
    }
    __syncthreads();
    fft_passes<OX, 1, kFftRowThreads, true>(sb, 1, t, stw, s);
    if (!live) return;
    const float syz = fg.scale * fg.fy[y] * fg.fz[z] * (((y + z) & 1) ? -1.0f : 1.0f);
    float2* out = img + rr * fg.Nx;
    for (int ix = t; ix < fg.Nx; ix += kFftRowThreads) {
        const int x = ix + fg.offx;
        float2 v = sb[fft_slot<true>(x, 1)];
        const float f = syz * fg.fx[x] * ((x & 1) ? -1.0f : 1.0f);
        v.x *= f;
        v.y *= f;
        out[ix] = v;
    }
}

// Strided lines of a 3D grid: line batch bb (of nbat) starts at
// (bb / nsub + off_o) * s_outer + (bb % nsub + off_i) * s_inner and holds
// ncols adjacent lines whose elements are rs apart; elements outside [lo, hi)
// are zero on input (BAND_IN) or not written (BAND_OUT).
struct FftLines {
    int nbat, nsub, off_o, off_i, ncols;
    long long s_outer, s_inner, rs;
    int lo, hi;
};

// Transforms of length OL along strided lines, W adjacent lines per
// workgroup (W = 8 with the XCD pairing of fft_cols_kernel; the 3D passes
// take 16 or 32, whose runs are whole 128-byte lines already).
template <int OL, bool BAND_IN, bool BAND_OUT, int W>
__global__ __launch_bounds__(256) void fft_lines_kernel(float2* __restrict__ grid, FftLines fl,
                                                        const float2* __restrict__ tw, float s) {
    constexpr int TPS = 256 / W;
    __shared__ float2 sbuf[OL * W];
    __shared__ float2 stw[OL / 2];
    const int nxb = (fl.ncols + W - 1) / W;
    int bb, xb;
    if (!fft_col_block<W>((int)blockIdx.x, fl.nbat, nxb, bb, xb)) return;  // block-uniform, before any barrier
    load_twiddles<OL>(stw, tw);
    const int xl = threadIdx.x % W, yr = threadIdx.x / W;
    const int x = xb * W + xl;
    const bool xin = x < fl.ncols;
    float2* line = grid + (long long)(bb / fl.nsub + fl.off_o) * fl.s_outer +
                   (long long)(bb % fl.nsub + fl.off_i) * fl.s_inner + (xin ? x : 0);
    {
        static_assert(OL % TPS == 0, "line length must be a multiple of the threads per line");
        constexpr int kPer = OL / TPS;
        float2 v[kPer];
#pragma unroll
        for (int i = 0; i < kPer; i++) {
            const int k = yr + i * TPS;
            v[i] = (xin && (!BAND_IN || (k >= fl.lo && k < fl.hi))) ? line[k * fl.rs] : make_float2(0.0f, 0.0f);
        }
#pragma unroll
        for (int i = 0; i < kPer; i++) sbuf[(yr + i * TPS) * W + xl] = v[i];
    }
    __syncthreads();
    fft_passes<OL, 1, TPS, false>(&sbuf[xl], W, yr, stw, s);
    if (!xin) return;
    for (int k = yr; k < OL; k += TPS) {
        if (BAND_OUT && (k < fl.lo || k >= fl.hi)) continue;
        line[k * fl.rs] = sbuf[k * W + xl];
    }
}

}  // namespace gt
