// gt_kernels_d.h -- the generic path's kernels in double precision, for plans
// made with GT_DOUBLE (NFFT<cuNDArray, double, D>, cuNFFT.cu:219-227: STANDARD
// convolution only).  Same algorithms and layouts as their float forms in
// gt_kernels.h (per-sample C2NC gather, per-tile NC2C with LDS accumulation in
// sample order and an ordered reduce of split tiles, pad / crop with
// deapodization, centring sign and scale); positions, weights, samples and
// accumulators are double.  The sample-tile binning runs on float copies of the
// positions with the window radius widened by kBinMargin cells, so that every
// tile a double-precision window touches is listed (extra tiles only receive
// zero weights).
#pragma once

#include "gt_kernels.h"

namespace gt {

constexpr float kBinMargin = 1.0e-3f;

// p = (k + 1/2) N_os in double, with a float copy for the binning; out-of-range
// or non-finite input raises the flag.
__global__ void scale_traj_d_kernel(const double* __restrict__ traj, double* __restrict__ out,
                                    float* __restrict__ outf, long long n, int D, GridGeom g,
                                    int* __restrict__ flag) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int d = (int)(i % D);
    const double k = traj[i];
    if (!(k >= -0.5 && k <= 0.5)) flag[0] = 1;
    const double v = (k + 0.5) * g.nos_d[d];
    out[i] = v;
    outf[i] = (float)v;
}

template <int KW>
__device__ __forceinline__ void window1d_d(double p, const GridGeom& g, int d, bool sign, double (&w)[KW],
                                           int (&cell)[KW]) {
    const double r = g.radius_d;
    const int lo = (int)ceil(p - r);
    const int hi = (int)floor(p + r);
#pragma unroll
    for (int i = 0; i < KW; i++) {
        const int c = lo + i;
        const double u = fabs((double)c - p);
        const bool ok = (c <= hi) && !(u > r);
        double wt = ok ? kb1d(u, g.nos_d[d], g.one_over_W_d, g.beta_d[d]) : 0.0;
        const int cw = wrap_cell(c, g.nos[d]);
        if (sign && (cw & 1)) wt = -wt;
        w[i] = wt;
        cell[i] = cw;
    }
}

// C2NC: one lane per sample, coils in register chunks of CC.
template <int D, int KW, int CC>
__global__ __launch_bounds__(256) void c2nc_gather_d_kernel(const double2* __restrict__ grid,
                                                            double2* __restrict__ out, const double* __restrict__ p,
                                                            int M, int F, int C, GridGeom g, int sign,
                                                            const double* __restrict__ dcw, long long dcw_len,
                                                            int accumulate) {
    const long long nsamp = (long long)M * F;
    const long long j = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nsamp) return;
    const int f = (int)(j / M);
    double wx[KW], wy[KW], wz[KW], wv[KW];
    int cx[KW], cy[KW], cz[KW], cv[KW];
    window1d_d<KW>(p[j * D + 0], g, 0, sign != 0, wx, cx);
    if (D >= 2) window1d_d<KW>(p[j * D + 1], g, 1, sign != 0, wy, cy);
    if (D >= 3) window1d_d<KW>(p[j * D + 2], g, 2, sign != 0, wz, cz);
    if (D >= 4) window1d_d<KW>(p[j * D + 3], g, 3, sign != 0, wv, cv);
    const long long G = g.G;
    const long long FG = (long long)F * G;
    const double2* gb = grid + (long long)f * G;
    for (int c0 = 0; c0 < C; c0 += CC) {
        double ar[CC], ai[CC];
#pragma unroll
        for (int q = 0; q < CC; q++) {
            ar[q] = 0.0;
            ai[q] = 0.0;
        }
        const int nv = D >= 4 ? KW : 1;
        const int nz = D >= 3 ? KW : 1;
        const int ny = D >= 2 ? KW : 1;
        for (int iv = 0; iv < nv; iv++)
        for (int iz = 0; iz < nz; iz++) {
            for (int iy = 0; iy < ny; iy++) {
#pragma unroll
                for (int ix = 0; ix < KW; ix++) {
                    double w;
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
                    if (w != 0.0) {
#pragma unroll
                        for (int q = 0; q < CC; q++) {
                            if (c0 + q < C) {
                                const double2 v = gb[(long long)(c0 + q) * FG + cell];
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
                const long long o = (long long)(c0 + q) * nsamp + j;
                double2 v = make_double2(ar[q], ai[q]);
                if (dcw) {
                    const double s = dcw[o % dcw_len];
                    v.x *= s;
                    v.y *= s;
                }
                if (accumulate) {
                    const double2 old = out[o];
                    v.x += old.x;
                    v.y += old.y;
                }
                out[o] = v;
            }
        }
    }
}

// NC2C: one wavefront per work item (a tile of one frame, a chunk of coils)
// accumulating in LDS; the tile's samples are taken in ascending order, 64 at
// a time, so each cell's sum order is fixed.
template <int D, int KW>
__global__ __launch_bounds__(64) void nc2c_tile_d_kernel(const double2* __restrict__ samples,
                                                         double2* __restrict__ grid, double2* __restrict__ partials,
                                                         const double* __restrict__ p,
                                                         const uint32_t* __restrict__ tile_samples,
                                                         const WorkItem* __restrict__ items, int M, int F, int C,
                                                         int CC, GridGeom g, int sign, const double* __restrict__ dcw,
                                                         long long dcw_len, int accumulate) {
    extern __shared__ double ldsd[];
    const WorkItem it = items[blockIdx.x];
    if (it.partial == kNoopItem) return;
    const int c0 = blockIdx.y * CC;
    const int ncc = min(CC, C - c0);
    const int TC = g.tile_cells;
    double* acc_re = ldsd;
    double* acc_im = ldsd + CC * TC;
    for (int i = threadIdx.x; i < 2 * CC * TC; i += 64) ldsd[i] = 0.0;
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
        double wx[KW], wy[KW], wz[KW], wv[KW];
        int cx[KW], cy[KW], cz[KW], cv[KW];
        window1d_d<KW>(p[(long long)j * D + 0], g, 0, false, wx, cx);
        if (D >= 2) window1d_d<KW>(p[(long long)j * D + 1], g, 1, false, wy, cy);
        if (D >= 3) window1d_d<KW>(p[(long long)j * D + 2], g, 2, false, wz, cz);
        if (D >= 4) window1d_d<KW>(p[(long long)j * D + 3], g, 3, false, wv, cv);
#pragma unroll
        for (int i = 0; i < KW; i++) {
            int l = cx[i] - org[0];
            cx[i] = (active && wx[i] != 0.0 && l >= 0 && l < g.tile[0]) ? l : -1;
            if (D >= 2) {
                l = cy[i] - org[1];
                cy[i] = (wy[i] != 0.0 && l >= 0 && l < g.tile[1]) ? l : -1;
            }
            if (D >= 3) {
                l = cz[i] - org[2];
                cz[i] = (wz[i] != 0.0 && l >= 0 && l < g.tile[2]) ? l : -1;
            }
            if (D >= 4) {
                l = cv[i] - org[3];
                cv[i] = (wv[i] != 0.0 && l >= 0 && l < g.tile[3]) ? l : -1;
            }
        }
        for (int q = 0; q < ncc; q++) {
            const long long sidx = (long long)(c0 + q) * FM + j;
            double2 s = active ? samples[sidx] : make_double2(0.0, 0.0);
            if (dcw && active) {
                const double dw = dcw[sidx % dcw_len];
                s.x *= dw;
                s.y *= dw;
            }
            double* are = acc_re + q * TC;
            double* aim = acc_im + q * TC;
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
                        double w;
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
        double2 v = make_double2(acc_re[q * TC + lc], acc_im[q * TC + lc]);
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
        if (sign && (par & 1)) {
            v.x = -v.x;
            v.y = -v.y;
        }
        const long long cell =
            gc[0] + (long long)g.nos[0] * (gc[1] + (long long)g.nos[1] * (gc[2] + (long long)g.nos[2] * gc[3]));
        const long long o = ((long long)(c0 + q) * F + f) * G + cell;
        if (accumulate) {
            const double2 old = grid[o];
            v.x += old.x;
            v.y += old.y;
        }
        grid[o] = v;
    }
}

// Sum of split tiles' partials in slot order.
__global__ __launch_bounds__(256) void nc2c_reduce_d_kernel(const double2* __restrict__ partials,
                                                            double2* __restrict__ grid,
                                                            const RedItem* __restrict__ red, int F, int C,
                                                            GridGeom g, int sign, int accumulate) {
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
        double2 v = make_double2(0.0, 0.0);
        for (int s = 0; s < it.count; s++) {
            const double2 a = partials[((long long)(it.first + s) * C + q) * TC + lc0];
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
        if (sign && (par & 1)) {
            v.x = -v.x;
            v.y = -v.y;
        }
        const long long cell =
            gc[0] + (long long)g.nos[0] * (gc[1] + (long long)g.nos[1] * (gc[2] + (long long)g.nos[2] * gc[3]));
        const long long o = ((long long)q * F + f) * G + cell;
        if (accumulate) {
            const double2 old = grid[o];
            v.x += old.x;
            v.y += old.y;
        }
        grid[o] = v;
    }
}

struct PadCropD {
    int N[4];
    int off[4];
    const double* filt[4];
    double scale;
    int sign;
    int in_os;
};

// Pad (or copy) + deapodize + centring sign + scale: a 64-thread row per grid line.
__global__ __launch_bounds__(64 * kRowsPerBlock) void pad_deapod_d_kernel(const double2* __restrict__ img,
                                                                          double2* __restrict__ grid, long long nb,
                                                                          GridGeom g, PadCropD pc) {
    const long long R = (long long)blockIdx.x * kRowsPerBlock + threadIdx.y;
    const int Ox = g.nos[0], Oy = g.nos[1], Oz = g.nos[2], Ov = g.nos[3];
    if (R >= nb * Oy * Oz * Ov) return;
    const int y = (int)(R % Oy);
    const long long t = R / Oy;
    const int z = (int)(t % Oz);
    const long long t2 = t / Oz;
    const int u = (int)(t2 % Ov);
    const long long b = t2 / Ov;
    double2* out = grid + R * Ox;
    const int iy = pc.in_os ? y : y - pc.off[1], iz = pc.in_os ? z : z - pc.off[2];
    const int iv = pc.in_os ? u : u - pc.off[3];
    const int Nx = pc.in_os ? Ox : pc.N[0], Ny = pc.in_os ? Oy : pc.N[1], Nz = pc.in_os ? Oz : pc.N[2];
    const int Nv = pc.in_os ? Ov : pc.N[3];
    const int offx = pc.in_os ? 0 : pc.off[0];
    if (iy < 0 || iy >= Ny || iz < 0 || iz >= Nz || iv < 0 || iv >= Nv) {
        for (int x = threadIdx.x; x < Ox; x += 64) out[x] = make_double2(0.0, 0.0);
        return;
    }
    const double2* in = img + (((b * Nv + iv) * Nz + iz) * Ny + iy) * Nx;
    double sr = pc.scale;
    if (g.D >= 2) sr *= pc.filt[1][y];
    if (g.D >= 3) sr *= pc.filt[2][z];
    if (g.D >= 4) sr *= pc.filt[3][u];
    if (pc.sign && ((y + z + u) & 1)) sr = -sr;
    for (int x = threadIdx.x; x < Ox; x += 64) {
        const int ix = x - offx;
        double2 v = make_double2(0.0, 0.0);
        if (ix >= 0 && ix < Nx) {
            v = in[ix];
            const double s = sr * pc.filt[0][x] * ((pc.sign && (x & 1)) ? -1.0 : 1.0);
            v.x *= s;
            v.y *= s;
        }
        out[x] = v;
    }
}

// Crop (or copy) + deapodize + centring sign + scale: a 64-thread row per output line.
__global__ __launch_bounds__(64 * kRowsPerBlock) void crop_deapod_d_kernel(const double2* in, double2* out,
                                                                           long long nb, GridGeom g, PadCropD pc) {
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
    const double2* src = in + (((b * Ov + u) * Oz + z) * Oy + y) * (long long)Ox + offx;
    double2* dst = out + R * Nx;
    double sr = pc.scale;
    if (g.D >= 2) sr *= pc.filt[1][y];
    if (g.D >= 3) sr *= pc.filt[2][z];
    if (g.D >= 4) sr *= pc.filt[3][u];
    if (pc.sign && ((y + z + u) & 1)) sr = -sr;
    for (int ix = threadIdx.x; ix < Nx; ix += 64) {
        const int x = ix + offx;
        double2 v = src[ix];
        const double s = sr * pc.filt[0][x] * ((pc.sign && (x & 1)) ? -1.0 : 1.0);
        v.x *= s;
        v.y *= s;
        dst[ix] = v;
    }
}

// data *= scale, times (-1)^(x+y+z) when sign is set.
__global__ __launch_bounds__(256) void sign_scale_d_kernel(double2* __restrict__ data, long long nb, GridGeom g,
                                                           double scale, int sign) {
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
        const double s = (sign && (par & 1)) ? -scale : scale;
        double2 v = data[i];
        v.x *= s;
        v.y *= s;
        data[i] = v;
    }
}

}  // namespace gt
