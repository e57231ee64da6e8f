// gt_kernels_reg.h -- the 2D gridding kernels of the hot path (8 x 8 tiles).
//
// nc2c_reg2d_kernel (NFFT_H_convolve_kernel, ConvolverNC2C_standard.cuh:298-355)
//   One wavefront owns one 8 x 8 tile of the grid for 32 coils; each lane
//   (one coil, real or imaginary part) holds the whole tile in 64 registers.
//   A table entry carries the sample's weight for each tile column and tile
//   row (zero outside the window) and the mask of rows the window covers, so
//   a sample costs 8 register FMAs per covered row, with the row test on the
//   scalar unit.  The table streams through LDS in chunks of kNK entries with
//   coalesced loads, two chunks ahead of the gridding, and the samples of the
//   next chunk are in flight while the current one is gridded.  Each lane
//   sums its cells in table order: the result is bitwise repeatable.  The
//   tile leaves through LDS, transposed for coalesced stores.
//
// c2nc_multi2d_kernel (NFFT_convolve_kernel, ConvolverC2NC.cuh:260-307)
//   A workgroup of kWaves wavefronts stages one tile plus the kernel halo
//   (32 coils, centring sign applied) and the work item's table in LDS; each
//   wavefront then takes a contiguous run of the tile's home samples and
//   gathers each window with all of its reads in flight before the FMAs
//   (two samples per step: with 32 coils a lane holds one coil's (re, im)).
#pragma once

#include "gt_kernels_fast.h"

namespace gt {

constexpr int kRT = kTabTile;  // register tile edge
constexpr int kWaves = 4;      // wavefronts per C2NC workgroup

// Arguments of one problem's matrix-core 2D gridding (either direction): the
// kernels below take them as launch arguments, and the batched launches of
// many problems from a per-problem table.
struct MfGridArgs {
    const float* in;      // NC2C: samples [M, F, C]; C2NC: the grid
    float* out;           // NC2C: the grid; C2NC: samples
    float* partials;      // NC2C: partial tiles of split tiles
    const float* wt;      // NC2C: table weights; C2NC: per-sample weight rows
    const uint32_t* hd;   // NC2C: table heads; C2NC: the tile-sorted sample list
    const WorkItem* items;
    int M, F, C;
    GridGeom g;
    Layout L;
    int sign;
    const float* dcw;
    long long dcw_len;
    int accumulate;
    float scale;          // NC2C: the weight normaliser; C2NC: the output scale
};

// One unit of a batched gridding launch: problem, work item, coil chunk
// (prob < 0: nothing to do).
struct GridUnit {
    int prob, item, chunk, pad_;
};

// Wave-uniform value the compiler cannot prove uniform by itself.
__device__ __forceinline__ int wave_uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }

constexpr int kNK = 16;  // NC2C table entries per chunk moved through LDS
constexpr int kLoadCoils = kCoilsPerWave / (kLanes / kNK);  // coils per lane in a chunk's sample loads
constexpr int kXsRow = 2 * kCoilsPerWave + 2;  // floats per entry of the sample transpose (+2: banks)
// LDS floats of nc2c_reg2d_kernel: max(write-out buffer, loop buffers)
constexpr int kNC2CLoopFloats = 2 * kNK * 4 * 4 + 2 * 2 * kNK + kNK * kXsRow;
constexpr int kNC2CLdsFloats = ((kRT * kRT / 2 * kSubP > kNC2CLoopFloats ? kRT * kRT / 2 * kSubP : kNC2CLoopFloats) + 3) & ~3;
static_assert(kNK == 16 && kLanes / kNK * kLoadCoils == kCoilsPerWave, "sample-load lane mapping");

// Loads chunk entries [cb, cb + kNK) of an NC2C table (entries [begin, end),
// end > begin) into registers, one float4 of weights per lane and one head
// dword per lane (lanes 32..63 repeat lanes 0..31).  Entries at or past `end`
// read as j = 0 with no rows (a no-op).  The loads are unconditional (clamped
// addresses, values selected afterwards), so every path through the caller
// issues the same number of memory operations.
__device__ __forceinline__ void nc2c_chunk_load(const float* __restrict__ wt, const uint32_t* __restrict__ hd,
                                                uint32_t cb, uint32_t end, int lane, float4& w, uint32_t& h) {
    const uint32_t e = cb + (uint32_t)(lane >> 2);
    const float4 wv = reinterpret_cast<const float4*>(wt)[(size_t)min(e, end - 1) * 4 + (lane & 3)];
    const bool we = e < end;
    w = make_float4(we ? wv.x : 0.0f, we ? wv.y : 0.0f, we ? wv.z : 0.0f, we ? wv.w : 0.0f);
    const uint32_t eh = cb + (uint32_t)((lane & 31) >> 1);
    const uint32_t hv = hd[(size_t)min(eh, end - 1) * 2 + (lane & 1)];
    h = eh < end ? hv : 0u;
}

// DM: how dcw is applied (dcw_mode): 0 none, 1 per sample (scalar loads),
// 2 per coil element (loaded with the samples).
// G: coil groups per wave (1, 2 or 4) for 32 / G coils.  Group g keeps its own
// copy of the tile and takes table entries g, g + G, ... of every chunk (rows
// follow the union of the groups' row masks; a row outside an entry's window
// has weight 0); the groups' tiles are added in a fixed order at the end, so
// the result stays bitwise repeatable.
// wt's column weights carry the power of two csc (nc2c_table2d_kernel); the
// sums are written times oscale = 1 / csc.
template <int DM, int G>
__global__ __launch_bounds__(64, 3) void nc2c_reg2d_kernel(
    const float* __restrict__ samples, float* __restrict__ grid, float* __restrict__ partials,
    const float* __restrict__ wt, const uint32_t* __restrict__ hd, const WorkItem* __restrict__ items, int M,
    int F, int C, GridGeom g, Layout L, int sign, const float* __restrict__ dcw, long long dcw_len,
    int accumulate, float oscale) {
    static_assert(G == 1 || G == 2 || G == 4, "coil groups");
    constexpr int CW = kCoilsPerWave / G;  // coils per wave
    constexpr int LC = CW / 4;             // coils per lane in a chunk's sample loads
    constexpr int XR = 2 * CW + 2;         // floats per entry of the sample transpose (+2: banks)
    constexpr int EK = kNK / G;            // entries per group and chunk
    // LDS: the gridding loop's buffers and the write-out's transpose buffer are
    // never live together (the write-out starts with a barrier)
    __shared__ float4 smem[kNC2CLdsFloats / 4];
    float* xp = reinterpret_cast<float*>(smem);  // half-tile transpose buffer
    float4 (*sw)[kNK * 4] = reinterpret_cast<float4 (*)[kNK * 4]>(smem);  // weights of two table chunks
    uint32_t (*sh)[2 * kNK] = reinterpret_cast<uint32_t (*)[2 * kNK]>(smem + 2 * kNK * 4);  // heads (j, rows)
    float* xs = reinterpret_cast<float*>(smem + 2 * kNK * 4 + kNK);  // a chunk's samples [entry][coil][re, im]
    const int lane = threadIdx.x;
    const WorkItem it = items[blockIdx.x];
    if (it.partial == kNoopItem) return;
    const int c0 = blockIdx.y * CW;
    const long long FM = (long long)F * M;
    const uint32_t nch = (it.end - it.begin + kNK - 1) / kNK;
    // sample loads: lane = (entry kk, coil quarter cq), LC coils each, so one
    // load instruction reads 4 coils x 16 entries -- runs of consecutive
    // samples of a spoke are contiguous per coil
    const int kk = lane & (kNK - 1), cq = lane >> 4;
    // gridding: lane = (group, coil, re/im)
    const int grp = lane / (2 * CW), q = lane % (2 * CW);
    const float2* s2 = reinterpret_cast<const float2*>(samples);

    float acc[kRT * kRT];
#pragma unroll
    for (int i = 0; i < kRT * kRT; i++) acc[i] = 0.0f;

    // Pipeline, per chunk c: the samples of chunk c (in registers) move to the
    // LDS transpose, the table of chunk c + 1 moves to LDS buffer (c + 1) & 1,
    // the table of chunk c + 2 and the samples of chunk c + 1 are issued, and
    // chunk c is gridded from LDS while those loads are in flight.
    constexpr int kDV = DM == 2 ? LC : 1;  // dcw per loaded coil (DM 2) or per entry (DM 1)
    float4 wreg;
    uint32_t hreg;
    float2 sv[LC];
    float dv[kDV];
#pragma unroll
    for (int u = 0; u < LC; u++) sv[u] = make_float2(0.0f, 0.0f);
#pragma unroll
    for (int u = 0; u < kDV; u++) dv[u] = 1.0f;

    auto stage = [&](int b) {
        sw[b][lane] = wreg;
        if (lane < 2 * kNK) sh[b][lane] = hreg;
    };
    auto issue = [&](int b) {
        const uint32_t j = sh[b][2 * kk];
#pragma unroll
        for (int u = 0; u < LC; u++) {
            int c = c0 + cq * LC + u;
            c = c < C ? c : c0;  // coils past the last one repeat a valid coil, never used
            sv[u] = s2[(long long)c * FM + j];
            if constexpr (DM == 2) dv[u] = dcw[dcw_index(c, j, FM, dcw_len)];
        }
        if constexpr (DM == 1) dv[0] = dcw[j];
    };
    auto transpose = [&]() {
        // lane = (entry, coil quarter) -> [entry][coil][re, im]; the gridding
        // reads lane = (group, coil, re/im) of its entry
#pragma unroll
        for (int u = 0; u < LC; u++) {
            float2 v = sv[u];
            if constexpr (DM != 0) {
                const float d = dv[DM == 2 ? u : 0];
                v.x *= d;
                v.y *= d;
            }
            *reinterpret_cast<float2*>(&xs[kk * XR + 2 * (cq * LC + u)]) = v;
        }
    };
    auto grid_chunk = [&](int b) {
        const int hv = (int)sh[b][lane & 31];
        // entry k + 1's weights and sample are read while entry k is gridded
        int e = grp;  // this lane's entry of step k: k * G + grp
        float4 n0 = sw[b][4 * e], n1 = sw[b][4 * e + 1], n2 = sw[b][4 * e + 2], n3 = sw[b][4 * e + 3];
        float nsk = xs[e * XR + q];
#pragma unroll
        for (int k = 0; k < EK; k++) {
            const float4 x0 = n0, x1 = n1, y0 = n2, y1 = n3;
            const float sk = nsk;
            if (k + 1 < EK) {
                e = (k + 1) * G + grp;
                n0 = sw[b][4 * e];
                n1 = sw[b][4 * e + 1];
                n2 = sw[b][4 * e + 2];
                n3 = sw[b][4 * e + 3];
                nsk = xs[e * XR + q];
            }
            uint32_t rows = 0;  // rows any group's entry covers
#pragma unroll
            for (int g = 0; g < G; g++) rows |= (uint32_t)__builtin_amdgcn_readlane(hv, 2 * (k * G + g) + 1);
            const float wc[kRT] = {x0.x, x0.y, x0.z, x0.w, x1.x, x1.y, x1.z, x1.w};
            const float wr[kRT] = {y0.x, y0.y, y0.z, y0.w, y1.x, y1.y, y1.z, y1.w};
            const float s = sk;  // dcw already applied in the transpose
            // rows: bits 0-7 covered tile rows (a scalar test per row); all 8
            // columns are updated (zero weight outside the window): skipping
            // half-rows costs more in branches than the FMAs it saves
#pragma unroll
            for (int cy = 0; cy < kRT; cy++) {
                if (!((rows >> cy) & 1u)) continue;
                const float sy = s * wr[cy];
#pragma unroll
                for (int cx = 0; cx < kRT; cx++) acc[cy * kRT + cx] = fmaf(sy, wc[cx], acc[cy * kRT + cx]);
            }
        }
    };

    // The steps past the last chunk stage and issue no-op entries rather than
    // branch, so the compiler's memory-counter waits see one steady state.
    if (nch > 0) {
        nc2c_chunk_load(wt, hd, it.begin, it.end, lane, wreg, hreg);
        stage(0);
        nc2c_chunk_load(wt, hd, it.begin + kNK, it.end, lane, wreg, hreg);
        issue(0);
        for (uint32_t c = 0; c < nch; c++) {
            const int b = (int)(c & 1u);
            transpose();
            stage(b ^ 1);
            nc2c_chunk_load(wt, hd, it.begin + (c + 2) * kNK, it.end, lane, wreg, hreg);
            issue(b ^ 1);
            grid_chunk(b);
        }
    }

    // the groups' tiles, added in a fixed order: (g0 + g1) + (g2 + g3)
    if constexpr (G == 4) {
#pragma unroll
        for (int i = 0; i < kRT * kRT; i++) acc[i] += __shfl_xor(acc[i], 16);
    }
    if constexpr (G >= 2) {
#pragma unroll
        for (int i = 0; i < kRT * kRT; i++) acc[i] += __shfl_xor(acc[i], 32);
    }

    // write-out through LDS, half a tile (4 rows) at a time: lanes run along x
    // within (coil, y) rows for coalesced stores (group 0 lanes hold the sums)
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
        for (int r = lr; r < CW * (kRT / 2); r += 4) {
            const int cl = r >> 2, ly = half * (kRT / 2) + (r & 3);
            const int cc = c0 + cl;
            const int gy = oy + ly;
            if (cc >= C || gx >= g.nos[0] || gy >= g.nos[1]) continue;
            float v = xp[(((r & 3) * kRT) + lx) * kSubP + 2 * cl + comp] * oscale;  // a power of two: exact
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

// Sum of the partial tiles of split tiles, in segment order (4 interleaved
// accumulators combined in a fixed order, so the result is repeatable).  One
// thread per (coil, cell, re/im) element of a tile; grid (split tiles, C*128/256).
__device__ __forceinline__ void nc2c_reduce2d_run(const float* __restrict__ partials, float* __restrict__ grid,
                                                  const RedItem it, int F, int C, const GridGeom& g,
                                                  const Layout& L, int sign, int accumulate) {
    const int i = blockIdx.y * blockDim.x + threadIdx.x;
    constexpr int kTileFloats = kRT * kRT * 2;
    if (i >= C * kTileFloats) return;
    const int comp = i & 1, cell = (i >> 1) & (kRT * kRT - 1), cc = i / kTileFloats;
    const int f = (int)(it.key / (uint32_t)g.ntiles_total);
    const int t = (int)(it.key % (uint32_t)g.ntiles_total);
    const int gx = (t % g.ntiles[0]) * kRT + (cell & (kRT - 1));
    const int gy = (t / g.ntiles[0]) * kRT + cell / kRT;
    if (gx >= g.nos[0] || gy >= g.nos[1]) return;
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
    if (sign && ((gx + gy) & 1)) v = -v;
    const long long o = 2 * ((long long)gx + L.sy * gy + L.sb * ((long long)cc * F + f)) + comp;
    if (accumulate) v += grid[o];
    grid[o] = v;
}

__global__ __launch_bounds__(256) void nc2c_reduce2d_kernel(const float* __restrict__ partials,
                                                            float* __restrict__ grid,
                                                            const RedItem* __restrict__ red, int F, int C,
                                                            GridGeom g, Layout L, int sign, int accumulate) {
    nc2c_reduce2d_run(partials, grid, red[blockIdx.x], F, C, g, L, sign, accumulate);
}

// Many problems: block (u, y) sums split tile u of the list (unit.item indexes
// the problem's RedItem list) for coil elements y * 256 ...
__global__ __launch_bounds__(256) void nc2c_reduce2d_batch_kernel(const MfGridArgs* __restrict__ probs,
                                                                  const RedItem* const* __restrict__ reds,
                                                                  const GridUnit* __restrict__ units) {
    const GridUnit u = units[blockIdx.x];
    if (u.prob < 0) return;
    const MfGridArgs& a = probs[u.prob];
    if ((int)(blockIdx.y * blockDim.x) >= a.C * kRT * kRT * 2) return;
    nc2c_reduce2d_run(a.partials, a.out, reds[u.prob][u.item], a.F, a.C, a.g, a.L, a.sign, a.accumulate);
}

// MODE 0: no dcw, no accumulation; 1: one dcw weight per sample, no
// accumulation; 2: anything else (per-element dcw, accumulation).
// G: coil groups (1, 2 or 4) for 32 / G coils.  G > 1: lane = (sample slot,
// coil, re/im), G samples per step; G = 1: lane = (sample slot, coil) reading
// (re, im) as one 8-byte word, 2 samples per step.  A per-sample dcw is
// loaded for each table entry while the tile is staged, and kept beside the
// table (dcw_e is unused here).
// LDS holds at most 3 workgroups per CU (3 waves per SIMD): the waves-per-EU
// range lets the scheduler keep a whole window's reads, and the staging's
// loads, in flight.
template <int KW, int MODE, int G>
__global__ __launch_bounds__(64 * kWaves) __attribute__((amdgpu_waves_per_eu(1, 3))) void c2nc_multi2d_kernel(
    const float* __restrict__ grid, float* __restrict__ out, const uint32_t* __restrict__ tab,
    const WorkItem* __restrict__ items, int M, int F, int C, GridGeom g, Layout L, int sign,
    const float* __restrict__ dcw, long long dcw_len, const float* __restrict__ dcw_e, int accumulate) {
    extern __shared__ float sg[];  // staged grid (c2nc_lds_floats), then the item's table (and dcw)
    static_assert(G == 1 || G == 2 || G == 4, "sample slots");
    constexpr int CW = kCoilsPerWave / G;  // coils per wave
    constexpr int SP = c2nc_cell_stride(G);  // LDS floats per staged cell
    constexpr int E0 = c2nc_extent(kRT, KW), E1 = c2nc_extent(kRT, KW);
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = wave_uniform(tid >> 6);
    const WorkItem it = items[blockIdx.x];
    if (it.partial == kNoopItem) return;
    const int c0 = blockIdx.y * CW;
    const int fr = (int)floorf(g.radius);
    const int f = (int)(it.key / (uint32_t)g.ntiles_total);
    const int t = (int)(it.key % (uint32_t)g.ntiles_total);
    const int sx = (t % g.ntiles[0]) * kRT - fr, sy0 = (t / g.ntiles[0]) * kRT - fr;
    const uint32_t ne = it.end - it.begin;  // at most kC2NCSeg
    uint4* st = reinterpret_cast<uint4*>(sg + c2nc_lds_floats(kRT, kRT, KW, G));
    float* sde = reinterpret_cast<float*>(st + kC2NCSeg * 4);  // per-entry dcw (G > 1, MODE 1)

    // the item's table entries, 16 dwords each (two uint4 per thread), loaded
    // together with the grid below and stored after it
    static_assert(kC2NCSeg * 4 <= 2 * 64 * kWaves, "two table uint4 per thread");
    constexpr int kTT = 64 * kWaves;
    const bool tlive = (uint32_t)tid < ne * 4, tlive2 = (uint32_t)(tid + kTT) < ne * 4;
    uint4 tv = make_uint4(0u, 0u, 0u, 0u), tv2 = make_uint4(0u, 0u, 0u, 0u);
    if (tlive) tv = reinterpret_cast<const uint4*>(tab)[(size_t)it.begin * 4 + tid];
    if (tlive2) tv2 = reinterpret_cast<const uint4*>(tab)[(size_t)it.begin * 4 + tid + kTT];
    // a per-sample dcw is staged per entry (lanes of a wave serve several
    // samples at once in every mode): the thread holding an entry's first
    // table word (its sample index) loads the entry's weight, in flight with
    // the grid loads below
    constexpr bool kDcwE = MODE == 1;
    const bool head = (tid & 3) == 0;
    float dj = 1.0f, dj2 = 1.0f;
    if (kDcwE && head && tlive) dj = dcw[tv.x];
    if (kDcwE && head && tlive2) dj2 = dcw[tv2.x];

    // tile + halo: a thread owns one float (cell ex, re or im) of a staged row
    // and walks the E1 rows for 4 coils; the row wrap and centring sign are
    // worked out once per row.  All E1 x 4 loads are issued before the first
    // LDS store, so the staging pays the memory latency once.
    {
        constexpr int kU = CW / 8;
        const int q = tid & 31, cg = tid >> 5;  // 8 coil groups of 32 threads
        const bool live = q < 2 * E0;
        const int ex = live ? q >> 1 : 0, comp = q & 1;
        int gx = sx + ex;
        gx = gx < 0 ? gx + g.nos[0] : (gx >= g.nos[0] ? gx - g.nos[0] : gx);
        const long long cstride = 2 * L.sb * (long long)F;  // next coil
        const float* gcol = grid + 2 * (long long)gx + comp + 2 * L.sb * (long long)f;
        float v[E1][kU];
#pragma unroll
        for (int ey = 0; ey < E1; ey++) {
            int gy = sy0 + ey;
            gy = gy < 0 ? gy + g.nos[1] : (gy >= g.nos[1] ? gy - g.nos[1] : gy);
            const float* prow = gcol + 2 * L.sy * (long long)gy;
#pragma unroll
            for (int u = 0; u < kU; u++) {
                const int cl = cg + 8 * u;
                v[ey][u] = (live && c0 + cl < C) ? prow[(long long)(c0 + cl) * cstride] : 0.0f;
            }
        }
        if (live) {
            float* dcol = sg + ex * SP + comp;
#pragma unroll
            for (int ey = 0; ey < E1; ey++) {
                const int gy = sy0 + ey;  // parity of the wrapped row equals that of gy (even N_os)
                const float sgn = (sign && ((gx + gy) & 1)) ? -1.0f : 1.0f;
#pragma unroll
                for (int u = 0; u < kU; u++) dcol[ey * E0 * SP + 2 * (cg + 8 * u)] = v[ey][u] * sgn;
            }
        }
        for (int i = tid; i < (E0 + 1) * SP; i += 64 * kWaves) sg[E0 * E1 * SP + i] = 0.0f;
    }
    if (tlive) st[tid] = tv;
    if (tlive2) st[tid + kTT] = tv2;
    if (kDcwE && head && tlive) sde[tid >> 2] = dj;
    if (kDcwE && head && tlive2) sde[(tid + kTT) >> 2] = dj2;
    __syncthreads();

    if constexpr (G == 1) {
        // 32 coils: lane = (slot of 2, coil) reads a coil's (re, im) as one
        // 8-byte word, so a wave gathers two samples per step with half the LDS
        // instructions of one lane per component; lanes past the last coil
        // mirror coil c0 (same values, same output address)
        const int slot = lane >> 5, coil = c0 + (lane & 31);
        const bool on = coil < C;
        const int lc = on ? coil : c0;
        const long long FM = (long long)F * M;
        const int dm = dcw_mode(dcw, dcw_len, FM);
        float2* ol = reinterpret_cast<float2*>(out) + (long long)lc * FM;
        constexpr int SP2 = SP / 2;
        const float2* sl = reinterpret_cast<const float2*>(sg) + (on ? (lane & 31) : 0);
        const uint32_t q0 = (ne * (uint32_t)wave) / kWaves, q1 = (ne * (uint32_t)(wave + 1)) / kWaves;
        auto gather2 = [&](uint32_t k, uint32_t& j) -> float2 {
            const uint4 h0 = st[4 * k], h1 = st[4 * k + 1], h2 = st[4 * k + 2], h3 = st[4 * k + 3];
            j = h0.x;
            const uint32_t pk = h0.y;
            const int bx = (int)(int8_t)(pk & 0xff), by = (int)(int8_t)((pk >> 8) & 0xff);
            const float wx[kTabKW] = {__uint_as_float(h0.z), __uint_as_float(h0.w), __uint_as_float(h1.x),
                                      __uint_as_float(h1.y), __uint_as_float(h1.z), __uint_as_float(h1.w),
                                      __uint_as_float(h2.x)};
            const float wy[kTabKW] = {__uint_as_float(h2.y), __uint_as_float(h2.z), __uint_as_float(h2.w),
                                      __uint_as_float(h3.x), __uint_as_float(h3.y), __uint_as_float(h3.z),
                                      __uint_as_float(h3.w)};
            const float2* base = sl + (by * E0 + bx) * SP2;
            float2 v[KW][KW];
#pragma unroll
            for (int iy = 0; iy < KW; iy++)
#pragma unroll
                for (int ix = 0; ix < KW; ix++) v[iy][ix] = base[(iy * E0 + ix) * SP2];
            float2 a = make_float2(0.0f, 0.0f);
#pragma unroll
            for (int iy = 0; iy < KW; iy++) {
                float rx = 0.0f, ry = 0.0f;
#pragma unroll
                for (int ix = 0; ix < KW; ix++) {
                    rx += wx[ix] * v[iy][ix].x;
                    ry += wx[ix] * v[iy][ix].y;
                }
                a.x += wy[iy] * rx;
                a.y += wy[iy] * ry;
            }
            return a;
        };
        for (uint32_t k = q0; k < q1; k += 2) {
            const uint32_t e = min(k + (uint32_t)slot, q1 - 1);  // a short tail repeats the last entry
            uint32_t j;
            float2 a = gather2(e, j);
            if constexpr (MODE == 1) {
                const float d = sde[e];
                a.x *= d;
                a.y *= d;
            } else if constexpr (MODE == 2) {
                float d = 1.0f;
                if (dm == 1) d = dcw[j];
                else if (dm == 2) d = dcw[dcw_index(lc, j, FM, dcw_len)];
                a.x *= d;
                a.y *= d;
                if (accumulate) {
                    const float2 o = ol[j];
                    a.x += o.x;
                    a.y += o.y;
                }
            }
            ol[j] = a;
        }
        return;
    }

    // lane = (slot, coil, re/im); lanes past the last coil mirror coil c0 (its
    // staged values, its output address): they store the same value to the
    // same place, so no lane masking
    const int slot = lane / (2 * CW), ql = lane % (2 * CW);
    const int coil = c0 + (ql >> 1);
    const bool on = coil < C;
    const int lc = on ? coil : c0;
    const long long FM = (long long)F * M;
    const int dm = dcw_mode(dcw, dcw_len, FM);
    float* ol = out + 2 * (long long)lc * FM + (ql & 1);
    const float* sl = sg + (on ? ql : (ql & 1));
    // wave w takes a contiguous quarter of the entries (consecutive samples of
    // a spoke share output cache lines), G at a time
    const uint32_t q0 = (ne * (uint32_t)wave) / kWaves, q1 = (ne * (uint32_t)(wave + 1)) / kWaves;
    // window sum of entry k; j is its sample index (wave-uniform for G = 1)
    auto gather = [&](uint32_t k, uint32_t& j) -> float {
        const uint4 h0 = st[4 * k], h1 = st[4 * k + 1], h2 = st[4 * k + 2], h3 = st[4 * k + 3];
        j = G == 1 ? (uint32_t)__builtin_amdgcn_readfirstlane((int)h0.x) : h0.x;
        const uint32_t pk = h0.y;
        const int bx = (int)(int8_t)(pk & 0xff), by = (int)(int8_t)((pk >> 8) & 0xff);
        const float wx[kTabKW] = {__uint_as_float(h0.z), __uint_as_float(h0.w), __uint_as_float(h1.x),
                                  __uint_as_float(h1.y), __uint_as_float(h1.z), __uint_as_float(h1.w),
                                  __uint_as_float(h2.x)};
        const float wy[kTabKW] = {__uint_as_float(h2.y), __uint_as_float(h2.z), __uint_as_float(h2.w),
                                  __uint_as_float(h3.x), __uint_as_float(h3.y), __uint_as_float(h3.z),
                                  __uint_as_float(h3.w)};
        // every slot of the window is read (slots outside it have weight 0
        // and stay inside the staged tile and its zeroed pad), all reads first
        const float* base = sl + (by * E0 + bx) * SP;
        float v[KW][KW];
#pragma unroll
        for (int iy = 0; iy < KW; iy++)
#pragma unroll
            for (int ix = 0; ix < KW; ix++) v[iy][ix] = base[(iy * E0 + ix) * SP];
        float a = 0.0f;
#pragma unroll
        for (int iy = 0; iy < KW; iy++) {
            float r = 0.0f;
#pragma unroll
            for (int ix = 0; ix < KW; ix++) r += wx[ix] * v[iy][ix];
            a += wy[iy] * r;
        }
        return a;
    };
    if constexpr (MODE < 2) {
        // no vector loads in the loop (a per-sample dcw is a scalar load, or
        // staged), so no wait ever covers the stores already issued
        for (uint32_t k = q0; k < q1; k += G) {
            const uint32_t e = G == 1 ? k : min(k + (uint32_t)slot, q1 - 1);  // a short tail repeats the last entry
            uint32_t j;
            float a = gather(e, j);
            if constexpr (MODE == 1) a *= G == 1 ? dcw[j] : sde[e];
            ol[2 * (long long)j] = a;
        }
    } else {
        for (uint32_t k = q0; k < q1; k += G) {
            const uint32_t e = G == 1 ? k : min(k + (uint32_t)slot, q1 - 1);
            uint32_t j;
            float a = gather(e, j);
            if (dm == 1) a *= dcw[j];
            else if (dm == 2) a *= dcw[dcw_index(lc, j, FM, dcw_len)];
            const long long o = 2 * (long long)j;
            if (accumulate) a += ol[o];
            ol[o] = a;
        }
    }
}

// ---------------------------------------------------------------------------
// nc2c_mfma2d_kernel: the 2D NC2C of 16 or 32 coils per wave (CW) on the
// matrix cores.
//
// A tile is a product of two matrices: grid[q][cell] = sum over entries k of
// s[q][k] * w[k][cell], with q = (coil, re/im) and w[k][cell] = wr_k[y] * wc_k[x].
// Each chunk of 16 table entries is one K = 16 step of
// v_mfma_f32_32x32x16_f16 on 2 CW / 32 x 2 blocks of 32 x 32 (rows (re / im,
// coil) x upper / lower half tile).  FP32 operands are split into fp16 pairs, x = hi + lo with
// hi = fp16(x) and lo = fp16(x - hi), and each block takes the three products
// hi*hi + hi*lo + lo*hi: every fp16 x fp16 product is exact in the FP32
// accumulator, and the dropped lo*lo term and the two roundings leave a
// relative error of about 2^-22 per product.  fp16 has a narrow exponent
// range, so the weights are normalised by a host power of two (wscale: the
// largest product of the two 1-D kernels becomes <= 1) and the samples by a
// power of two R kept per wave: R is set from the first chunk with nonzero
// samples to put them below 2^12, and lowered (the accumulators rescaled
// exactly) when a chunk would exceed 2^15.  Results are multiplied by
// 1 / (R * wscale) on the way out.  Entry order, chunk order and the
// hardware's product order are fixed: the result is bitwise repeatable.
//
// Operand lane layouts (32x32x16 f16): A holds row l % 32 at k = 8 (l / 32)
// + i; B holds column l % 32 at the same k; accumulator register r of lane l
// is row 8 (r / 4) + 4 (l / 32) + r % 4, column l % 32.
// ---------------------------------------------------------------------------
typedef _Float16 mf_half8 __attribute__((ext_vector_type(8)));
typedef float mf_float8 __attribute__((ext_vector_type(8)));
typedef float mf_float16 __attribute__((ext_vector_type(16)));
typedef float mf_float4 __attribute__((ext_vector_type(4)));

// CW: coils per wave (32: two 32-row blocks, real and imaginary parts; 16: one
// block, rows = (re / im, coil)).
template <int CW>
struct MfLds {
    static constexpr int WS = kNK + 4;         // floats per row of a transposed table chunk
    static constexpr int XS = kNK + 4;         // floats per sample row (re or im of one coil)
    static constexpr int Tab = 2 * 16 * WS;    // two chunks: rows 0-7 column weights, 8-15 row weights
    static constexpr int Heads = 2 * 2 * kNK;
    static constexpr int Loop = Tab + Heads + 2 * CW * XS;
    static constexpr int Out = kRT * kRT / 2 * kSubP;
    static constexpr int Floats = ((Out > Loop ? Out : Loop) + 3) & ~3;
};

__device__ __forceinline__ void mf_split(const mf_float8& x, mf_half8& hi, mf_half8& lo) {
    hi = __builtin_convertvector(x, mf_half8);
    lo = __builtin_convertvector(x - __builtin_convertvector(hi, mf_float8), mf_half8);
}

__device__ __forceinline__ float mf_wave_max(float v) {
    // v >= 0: the bit patterns order as the values do
    uint32_t b = __float_as_uint(v);
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) b = max(b, (uint32_t)__shfl_xor((int)b, o));
    return __uint_as_float(b);
}

__device__ __forceinline__ float mf_absmax(const mf_float8& x) {
    return fmaxf(fmaxf(fmaxf(fabsf(x[0]), fabsf(x[1])), fmaxf(fabsf(x[2]), fabsf(x[3]))),
                 fmaxf(fmaxf(fabsf(x[4]), fabsf(x[5])), fmaxf(fabsf(x[6]), fabsf(x[7]))));
}

template <int DM, int CW>
__device__ __forceinline__ void nc2c_mfma2d_run(const MfGridArgs& a, int item, int chunk) {
    const float* __restrict__ samples = a.in;
    float* __restrict__ grid = a.out;
    float* __restrict__ partials = a.partials;
    const float* __restrict__ wt = a.wt;
    const uint32_t* __restrict__ hd = a.hd;
    const WorkItem* __restrict__ items = a.items;
    const int M = a.M, F = a.F, C = a.C;
    const GridGeom& g = a.g;
    const Layout L = a.L;
    const int sign = a.sign, accumulate = a.accumulate;
    const float* __restrict__ dcw = a.dcw;
    const long long dcw_len = a.dcw_len;
    const float wscale = a.scale;
    static_assert(CW == 32 || CW == 16, "coils per wave");
    constexpr int LC = CW / 4;       // coils per lane in a chunk's sample loads
    constexpr int MB = 2 * CW / 32;  // 32-row blocks of (coil, re / im)
    using Lds = MfLds<CW>;
    constexpr int WS = Lds::WS, XS = Lds::XS;
    __shared__ float4 smem[Lds::Floats / 4];
    float* xp = reinterpret_cast<float*>(smem);                 // write-out transpose
    float* wT = reinterpret_cast<float*>(smem);                 // [2][16][WS]
    uint32_t* sh = reinterpret_cast<uint32_t*>(wT + Lds::Tab);  // [2][2 kNK] heads
    float* xs = reinterpret_cast<float*>(sh + Lds::Heads);      // [re / im][coil][XS] per entry
    const int lane = threadIdx.x;
    const WorkItem it = items[item];
    if (it.partial == kNoopItem) return;
    const int c0 = chunk * CW;
    const long long FM = (long long)F * M;
    const uint32_t nch = (it.end - it.begin + kNK - 1) / kNK;
    const int kk = lane & (kNK - 1), cq = lane >> 4;  // sample loads: (entry, coil quarter)
    const int col = lane & 31, hk = lane >> 5;         // MFMA: row / column l % 32, k half
    const float2* s2 = reinterpret_cast<const float2*>(samples);

    mf_float16 acc[MB][2];
#pragma unroll
    for (int a = 0; a < MB; a++)
#pragma unroll
        for (int b = 0; b < 2; b++)
#pragma unroll
            for (int r = 0; r < 16; r++) acc[a][b][r] = 0.0f;
    float R = 0.0f;  // sample scale (0: not set yet)

    constexpr int kDV = DM == 2 ? LC : 1;
    float4 wreg;
    uint32_t hreg;
    float2 sv[LC];
    float dv[kDV];
#pragma unroll
    for (int u = 0; u < LC; u++) sv[u] = make_float2(0.0f, 0.0f);
#pragma unroll
    for (int u = 0; u < kDV; u++) dv[u] = 1.0f;

    auto stage = [&](int b) {
        // lane = (entry e, part p): part 0 / 1 column weights 0-3 / 4-7, 2 / 3 row weights
        const int e = lane >> 2, pr = lane & 3;
        float* dst = wT + b * 16 * WS + 4 * pr * WS + e;
        dst[0] = wreg.x;
        dst[WS] = wreg.y;
        dst[2 * WS] = wreg.z;
        dst[3 * WS] = wreg.w;
        if (lane < 2 * kNK) sh[b * 2 * kNK + lane] = hreg;
    };
    auto issue = [&](int b) {
        const uint32_t j = sh[b * 2 * kNK + 2 * kk];
#pragma unroll
        for (int u = 0; u < LC; u++) {
            int c = c0 + cq * LC + u;
            c = c < C ? c : c0;  // coils past the last one repeat a valid coil, never stored
            sv[u] = s2[(long long)c * FM + j];
            if constexpr (DM == 2) dv[u] = dcw[dcw_index(c, j, FM, dcw_len)];
        }
        if constexpr (DM == 1) dv[0] = dcw[j];
    };
    auto transpose = [&]() {
#pragma unroll
        for (int u = 0; u < LC; u++) {
            float2 v = sv[u];
            if constexpr (DM != 0) {
                const float d = dv[DM == 2 ? u : 0];
                v.x *= d;
                v.y *= d;
            }
            const int cl = cq * LC + u;
            xs[cl * XS + kk] = v.x;
            xs[(CW + cl) * XS + kk] = v.y;
        }
    };
    auto grid_chunk = [&](int b) {
        // A: rows a * 32 + col (re / im, coil), entries 8 hk .. 8 hk + 7
        mf_float8 x[MB];
        const float Rt = R == 0.0f ? 1.0f : R;
        float m = 0.0f;
#pragma unroll
        for (int a = 0; a < MB; a++) {
            const float4* xa = reinterpret_cast<const float4*>(xs + (a * 32 + col) * XS + 8 * hk);
            const float4 a0 = xa[0], a1 = xa[1];
            x[a] = mf_float8{a0.x, a0.y, a0.z, a0.w, a1.x, a1.y, a1.z, a1.w} * Rt;
            m = fmaxf(m, mf_absmax(x[a]));
        }
        if (__builtin_amdgcn_ballot_w64(m > 32768.0f || (R == 0.0f && m > 0.0f))) {
            const float mx = mf_wave_max(m);  // of x = samples * Rt
            int e;
            frexpf(mx, &e);                   // mx < 2^e
            e = min(max(e, -100), 140);       // keep R and 1 / R finite (denormal or huge samples)
            const float f = ldexpf(1.0f, 12 - e);  // x * f < 2^12; R * f is the new scale
            if (R != 0.0f) {
#pragma unroll
                for (int a = 0; a < MB; a++)
#pragma unroll
                    for (int h = 0; h < 2; h++) acc[a][h] *= f;  // powers of two: exact
            }
#pragma unroll
            for (int a = 0; a < MB; a++) x[a] *= f;
            R = Rt * f;
        }
        mf_half8 xh[MB], xl[MB];
#pragma unroll
        for (int a = 0; a < MB; a++) mf_split(x[a], xh[a], xl[a]);
        // B: column weight of this lane's cell x, row weights of its y in both half tiles
        const float* wb = wT + b * 16 * WS + 8 * hk;
        const int cx = col & 7, cy = col >> 3;
        const float4* pc = reinterpret_cast<const float4*>(wb + cx * WS);
        const float4* p0 = reinterpret_cast<const float4*>(wb + (8 + cy) * WS);
        const float4* p1 = reinterpret_cast<const float4*>(wb + (12 + cy) * WS);
        const float4 c_0 = pc[0], c_1 = pc[1], r00 = p0[0], r01 = p0[1], r10 = p1[0], r11 = p1[1];
        // the table's column weights carry wscale already
        const mf_float8 wc = mf_float8{c_0.x, c_0.y, c_0.z, c_0.w, c_1.x, c_1.y, c_1.z, c_1.w};
        const mf_float8 w0 = wc * mf_float8{r00.x, r00.y, r00.z, r00.w, r01.x, r01.y, r01.z, r01.w};
        const mf_float8 w1 = wc * mf_float8{r10.x, r10.y, r10.z, r10.w, r11.x, r11.y, r11.z, r11.w};
        mf_half8 w0h, w0l, w1h, w1l;
        mf_split(w0, w0h, w0l);
        mf_split(w1, w1h, w1l);
#pragma unroll
        for (int a = 0; a < MB; a++) {
            acc[a][0] = __builtin_amdgcn_mfma_f32_32x32x16_f16(xl[a], w0h, acc[a][0], 0, 0, 0);
            acc[a][1] = __builtin_amdgcn_mfma_f32_32x32x16_f16(xl[a], w1h, acc[a][1], 0, 0, 0);
        }
#pragma unroll
        for (int a = 0; a < MB; a++) {
            acc[a][0] = __builtin_amdgcn_mfma_f32_32x32x16_f16(xh[a], w0l, acc[a][0], 0, 0, 0);
            acc[a][1] = __builtin_amdgcn_mfma_f32_32x32x16_f16(xh[a], w1l, acc[a][1], 0, 0, 0);
        }
#pragma unroll
        for (int a = 0; a < MB; a++) {
            acc[a][0] = __builtin_amdgcn_mfma_f32_32x32x16_f16(xh[a], w0h, acc[a][0], 0, 0, 0);
            acc[a][1] = __builtin_amdgcn_mfma_f32_32x32x16_f16(xh[a], w1h, acc[a][1], 0, 0, 0);
        }
    };

    if (nch > 0) {
        nc2c_chunk_load(wt, hd, it.begin, it.end, lane, wreg, hreg);
        stage(0);
        nc2c_chunk_load(wt, hd, it.begin + kNK, it.end, lane, wreg, hreg);
        issue(0);
        for (uint32_t c = 0; c < nch; c++) {
            const int b = (int)(c & 1u);
            transpose();
            stage(b ^ 1);
            nc2c_chunk_load(wt, hd, it.begin + (c + 2) * kNK, it.end, lane, wreg, hreg);
            issue(b ^ 1);
            grid_chunk(b);
        }
    }

    // write-out through LDS, half a tile (4 rows) at a time, as nc2c_reg2d_kernel
    const float outs = R > 0.0f ? 1.0f / (R * wscale) : 0.0f;  // powers of two: exact
    const int f = (int)(it.key / (uint32_t)g.ntiles_total);
    const int t = (int)(it.key % (uint32_t)g.ntiles_total);
    const int ox = (t % g.ntiles[0]) * kRT, oy = (t / g.ntiles[0]) * kRT;
    const int lr = lane >> 4, li = lane & 15;
    const int lx = li >> 1, comp = li & 1;
    const int gx = ox + lx;
#pragma unroll
    for (int half = 0; half < 2; half++) {
        __syncthreads();
#pragma unroll
        for (int a = 0; a < MB; a++)
#pragma unroll
            for (int r = 0; r < 16; r++) {
                const int row = a * 32 + 8 * (r >> 2) + 4 * hk + (r & 3);  // (re / im, coil)
                xp[col * kSubP + 2 * (row % CW) + row / CW] = acc[a][half][r] * outs;
            }
        __syncthreads();
        for (int r = lr; r < CW * (kRT / 2); r += 4) {
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

// Block b = (item / 8) * 8 nchunks + chunk * 8 + item % 8: the coil chunks of an
// item run on its XCD (b mod 8 = item mod 8), 8 blocks apart, and share its L2.
template <int DM, int CW>
__global__ __launch_bounds__(64, CW == 32 ? 2 : 3) void nc2c_mfma2d_kernel(
    const float* __restrict__ samples, float* __restrict__ grid, float* __restrict__ partials,
    const float* __restrict__ wt, const uint32_t* __restrict__ hd, const WorkItem* __restrict__ items, int M,
    int F, int C, GridGeom g, Layout L, int sign, const float* __restrict__ dcw, long long dcw_len,
    int accumulate, float wscale, int n_items, int nchunks) {
    const int bx = (int)blockIdx.x, grp = 8 * nchunks;
    const int item = (bx / grp) * 8 + (bx & 7), chunk = (bx / 8) % nchunks;
    if (item >= n_items) return;
    const MfGridArgs a{samples, grid, partials, wt, hd, items, M, F, C, g, L, sign, dcw, dcw_len, accumulate, wscale};
    nc2c_mfma2d_run<DM, CW>(a, item, chunk);
}

// Many problems in one launch: block b runs unit b of the list, with the
// arguments of the unit's problem.
template <int DM, int CW>
__global__ __launch_bounds__(64, CW == 32 ? 2 : 3) void nc2c_mfma2d_batch_kernel(
    const MfGridArgs* __restrict__ probs, const GridUnit* __restrict__ units) {
    const GridUnit u = units[blockIdx.x];
    if (u.prob < 0) return;
    const MfGridArgs a = probs[u.prob];
    nc2c_mfma2d_run<DM, CW>(a, u.item, u.chunk);
}

// ---------------------------------------------------------------------------
// c2nc_mfma2d_kernel: the 2D C2NC of 16 or 32 coils per workgroup (CW) on
// the matrix cores.
//
// The samples of a home tile are a product: s[q][n] = sum over staged cells k
// of grid[q][k] * w[k][n], q = (coil, re/im), n = sample, w[k][n] = wy_n[y] *
// wx_n[x].  Home tiles are c2nc_mfma_tile_x(KW) x 8 cells, so that the tile +
// halo is 16 columns (one K = 16 step per staged row) by E = c2nc_extent(8, KW)
// rows.  A workgroup of four waves stages it once, scaled by a power of two R and
// split into fp16 hi / lo planes in LDS; each wave then takes 32 samples and
// walks the E staged rows as K = 16 steps of v_mfma_f32_32x32x16_f16 on
// 2 CW / 32 blocks of 32 rows (16 coils each, rows 2 coil + re/im), with the three
// products hi*hi + hi*lo + lo*hi as in nc2c_mfma2d_kernel.  The B operand is
// the sample's weight row product, from a table that holds each sample's
// weights at staged coordinates (c2nc_wtable2d_kernel).  Out: s * oscale / R
// (oscale undoes the table's power-of-two weight scales), then dcw.
// ---------------------------------------------------------------------------
__host__ __device__ constexpr int c2nc_mfma_row(int KW) { return c2nc_extent(kRT, KW) * 16 + 8; }  // halves per A row
// Home tile width (x) of the matrix-core C2NC: its halo fills a staged row of
// 16 cells (11 cells at windows of 5 or 6 cells, 9 at 7, 13 at 3 or 4); the
// tiles are kRT rows high.
__host__ __device__ constexpr int c2nc_mfma_tile_x(int KW) { return 16 - c2nc_extent(0, KW); }
__host__ __device__ constexpr int c2nc_mfma_lds_bytes(int KW, int CW) {
    return 2 * (2 * CW) * c2nc_mfma_row(KW) * 2 + 64;
}

// CW: coils per workgroup (32 or 16: two or one 32-row blocks).
template <int KW, int CW>
__device__ __forceinline__ void c2nc_mfma2d_run(const MfGridArgs& a, int item, int chunk) {
    const float* __restrict__ grid = a.in;
    float* __restrict__ out = a.out;
    const float* __restrict__ wtab = a.wt;
    const uint32_t* __restrict__ list = a.hd;
    const WorkItem* __restrict__ items = a.items;
    const int M = a.M, F = a.F, C = a.C;
    const GridGeom& g = a.g;
    const Layout L = a.L;
    const int sign = a.sign, accumulate = a.accumulate;
    const float* __restrict__ dcw = a.dcw;
    const long long dcw_len = a.dcw_len;
    const float oscale = a.scale;
    constexpr int E = c2nc_extent(kRT, KW);  // staged rows
    constexpr int AS = c2nc_mfma_row(KW);
    constexpr int MB = 2 * CW / 32;  // 32-row blocks
    static_assert(E <= 16 && (CW == 32 || CW == 16), "staged rows of 16 cells, 16 or 32 coils");
    extern __shared__ float4 c2m_smem[];
    _Float16* Ahi = reinterpret_cast<_Float16*>(c2m_smem);
    _Float16* Alo = Ahi + 2 * CW * AS;
    float* red = reinterpret_cast<float*>(Alo + 2 * CW * AS);  // four wave maxima
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const WorkItem it = items[item];
    if (it.partial == kNoopItem) return;
    const int c0 = chunk * CW;
    const int fr = (int)floorf(g.radius);
    const int f = (int)(it.key / (uint32_t)g.ntiles_total);
    const int t = (int)(it.key % (uint32_t)g.ntiles_total);
    const int sx = (t % g.ntiles[0]) * g.tile[0] - fr, sy0 = (t / g.ntiles[0]) * kRT - fr;
    const uint32_t ne = it.end - it.begin;  // at most 4 x 32
    const int EX = c2nc_extent(g.tile[0], KW);  // staged columns (16 with c2nc_mfma_tile_x)

    // this lane's sample (MFMA column) and its weight rows, loaded before the
    // staging so that the two latencies overlap
    const int col = lane & 31, hk = lane >> 5;
    const uint32_t e = (uint32_t)(wave * 32 + col);
    const bool on = e < ne;
    const uint32_t j = list[it.begin + (on ? e : 0u)];
    const float4* wr = reinterpret_cast<const float4*>(wtab + (size_t)(it.begin + (on ? e : 0u)) * 32);
    const float4 xa = wr[2 * hk], xb = wr[2 * hk + 1];
    const float4 y0 = wr[4], y1 = wr[5], y2 = wr[6], y3 = wr[7];

    // staging: thread = (cell column ex, coil), each cell's (re, im) read as
    // one 8-byte word; columns past EX write zeros, so every staged row holds
    // 16 defined cells
    constexpr int kU = CW / 16;
    const int ex = tid & 15, cg = tid >> 4;
    const bool live = ex < EX;
    int gx = sx + ex;
    gx = gx < 0 ? gx + g.nos[0] : (gx >= g.nos[0] ? gx - g.nos[0] : gx);
    const long long cstride = L.sb * (long long)F;  // next coil, in cells
    const float2* gcol = reinterpret_cast<const float2*>(grid) + (long long)(live ? gx : 0) + L.sb * (long long)f;
    float2 v[E][kU];
    float mloc = 0.0f;
#pragma unroll
    for (int ey = 0; ey < E; ey++) {
        int gy = sy0 + ey;
        gy = gy < 0 ? gy + g.nos[1] : (gy >= g.nos[1] ? gy - g.nos[1] : gy);
        const float2* prow = gcol + L.sy * (long long)gy;
#pragma unroll
        for (int u = 0; u < kU; u++) {
            const int cl = cg + 16 * u;
            v[ey][u] = (live && c0 + cl < C) ? prow[(long long)(c0 + cl) * cstride] : make_float2(0.0f, 0.0f);
        }
    }
    const long long FM = (long long)F * M;
    const int dm = dcw_mode(dcw, dcw_len, FM);
    const float dj = dm == 1 ? dcw[j] : 1.0f;  // in flight with the staging
#pragma unroll
    for (int ey = 0; ey < E; ey++) {
        const int gy = sy0 + ey;  // parity of the wrapped row equals that of gy (even N_os)
        const float sgn = (sign && ((gx + gy) & 1)) ? -1.0f : 1.0f;
#pragma unroll
        for (int u = 0; u < kU; u++) {
            v[ey][u].x *= sgn;
            v[ey][u].y *= sgn;
            mloc = fmaxf(mloc, fmaxf(fabsf(v[ey][u].x), fabsf(v[ey][u].y)));
        }
    }
    const float wm = mf_wave_max(mloc);
    if (lane == 0) red[wave] = wm;
    __syncthreads();
    const float mx = fmaxf(fmaxf(red[0], red[1]), fmaxf(red[2], red[3]));
    int e2;
    frexpf(mx, &e2);               // mx < 2^e2
    e2 = min(max(e2, -100), 140);  // keep R and 1 / R finite
    const float R = mx > 0.0f ? ldexpf(1.0f, 12 - e2) : 1.0f;
#pragma unroll
    for (int ey = 0; ey < E; ey++)
#pragma unroll
        for (int u = 0; u < kU; u++) {
            const float xr = v[ey][u].x * R, xi = v[ey][u].y * R;
            const _Float16 hr = (_Float16)xr, hi = (_Float16)xi;
            const _Float16 lr = (_Float16)(xr - (float)hr), li = (_Float16)(xi - (float)hi);
            const int idx = (2 * (cg + 16 * u)) * AS + ey * 16 + ex;  // rows 2 coil (re), 2 coil + 1 (im)
            Ahi[idx] = hr;
            Alo[idx] = lr;
            Ahi[idx + AS] = hi;
            Alo[idx + AS] = li;
        }
    __syncthreads();

    if ((uint32_t)(wave * 32) >= ne) return;  // wave-uniform; no barrier follows
    const float z = on ? 1.0f : 0.0f;
    const mf_float8 wx = mf_float8{xa.x, xa.y, xa.z, xa.w, xb.x, xb.y, xb.z, xb.w} * z;
    const float wy[16] = {y0.x, y0.y, y0.z, y0.w, y1.x, y1.y, y1.z, y1.w,
                          y2.x, y2.y, y2.z, y2.w, y3.x, y3.y, y3.z, y3.w};
    mf_float16 acc[MB];
#pragma unroll
    for (int a = 0; a < MB; a++)
#pragma unroll
        for (int r = 0; r < 16; r++) acc[a][r] = 0.0f;
    const _Float16* ah = Ahi + col * AS + 8 * hk;
    const _Float16* al = Alo + col * AS + 8 * hk;
#pragma unroll
    for (int sy = 0; sy < E; sy++) {
        mf_half8 bh, bl;
        mf_split(wx * wy[sy], bh, bl);
        mf_half8 xh[MB], xl[MB];
#pragma unroll
        for (int a = 0; a < MB; a++) {
            xh[a] = *reinterpret_cast<const mf_half8*>(ah + a * 32 * AS + sy * 16);
            xl[a] = *reinterpret_cast<const mf_half8*>(al + a * 32 * AS + sy * 16);
        }
#pragma unroll
        for (int a = 0; a < MB; a++) acc[a] = __builtin_amdgcn_mfma_f32_32x32x16_f16(xl[a], bh, acc[a], 0, 0, 0);
#pragma unroll
        for (int a = 0; a < MB; a++) acc[a] = __builtin_amdgcn_mfma_f32_32x32x16_f16(xh[a], bl, acc[a], 0, 0, 0);
#pragma unroll
        for (int a = 0; a < MB; a++) acc[a] = __builtin_amdgcn_mfma_f32_32x32x16_f16(xh[a], bh, acc[a], 0, 0, 0);
    }
    if (!on) return;
    // accumulator register r of block a: row 32 a + 8 (r / 4) + 4 hk + r % 4 =
    // 2 coil + re/im, column = this lane's sample: (r, r + 1) for even r are
    // one coil's (re, im)
    const float d = oscale / R * dj;
    float2* o2 = reinterpret_cast<float2*>(out);
#pragma unroll
    for (int a = 0; a < MB; a++)
#pragma unroll
        for (int r = 0; r < 16; r += 2) {
            const int cl = (32 * a + 8 * (r >> 2) + 4 * hk + (r & 3)) >> 1;
            const int cc = c0 + cl;
            if (cc >= C) continue;
            float dd = d;
            if (dm == 2) dd *= dcw[dcw_index(cc, j, FM, dcw_len)];
            float2 val = make_float2(acc[a][r] * dd, acc[a][r + 1] * dd);
            float2* o = o2 + (long long)cc * FM + j;
            if (accumulate) {
                const float2 old = *o;
                val.x += old.x;
                val.y += old.y;
            }
            *o = val;
        }
}

// Block b = (item / 8) * 8 nchunks + chunk * 8 + item % 8, as the NC2C kernel.
template <int KW, int CW>
__global__ __launch_bounds__(256) void c2nc_mfma2d_kernel(
    const float* __restrict__ grid, float* __restrict__ out, const float* __restrict__ wtab,
    const uint32_t* __restrict__ list, const WorkItem* __restrict__ items, int M, int F, int C, GridGeom g,
    Layout L, int sign, const float* __restrict__ dcw, long long dcw_len, int accumulate, float oscale, int n_items,
    int nchunks) {
    const int bx = (int)blockIdx.x, grp = 8 * nchunks;
    const int item = (bx / grp) * 8 + (bx & 7), chunk = (bx / 8) % nchunks;
    if (item >= n_items) return;
    const MfGridArgs a{grid, out, nullptr, wtab, list, items, M, F, C, g, L, sign, dcw, dcw_len, accumulate, oscale};
    c2nc_mfma2d_run<KW, CW>(a, item, chunk);
}

template <int KW, int CW>
__global__ __launch_bounds__(256) void c2nc_mfma2d_batch_kernel(const MfGridArgs* __restrict__ probs,
                                                                const GridUnit* __restrict__ units) {
    const GridUnit u = units[blockIdx.x];
    if (u.prob < 0) return;
    const MfGridArgs a = probs[u.prob];
    c2nc_mfma2d_run<KW, CW>(a, u.item, u.chunk);
}

}  // namespace gt
