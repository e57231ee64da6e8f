// gt_cg.hip -- the non-Cartesian SENSE operator and its CG solver over the
// NFFT plan (include/gt_cgsense.h).
//
// SENSE (cuNonCartesianSenseOperator.cu:14-62, sense_utilities.cu:7-123):
// mult_M multiplies the image by every coil map and runs FORWARDS_C2NC on the
// [N..., F, C] stack; mult_MH runs BACKWARDS_NC2C and sums conj(csm) times the
// coil images.
//
// CG (cgSolver.h:219-390): the reference reads every dot product back to the
// host (cuBLAS cdotc) and forms alpha, beta there.  Here the dot products are
// FP64 block reductions finished on the device by one-block kernels that also
// form the step sizes and the termination flag, so an iteration issues no
// host round trip and a whole solve can be captured into a hipGraph.
// Reductions run in a fixed order: results repeat bit for bit, and the plain
// and batched solvers agree exactly.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <fstream>
#include <map>
#include <memory>
#include <mutex>
#include <numeric>
#include <string>
#include <utility>
#include <vector>

#include "../../include/gt_cgsense.h"
#include "gt_common.h"

using namespace gt_host;

namespace {

constexpr int kRedBlocks = 512;   // partial sums per FP64 reduction
constexpr int kRedThreads = 256;

// ---------------------------------------------------------------------------
// coil kernels (sense_utilities.cu:7-123)
// ---------------------------------------------------------------------------
// out[n, f, c] = in[n, f] * csm[n, c]
__global__ void csm_mult_kernel(const float2* __restrict__ in, float2* __restrict__ out,
                                const float2* __restrict__ csm, long long nimg, int F, int C) {
    const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const int f = blockIdx.y;
    if (n >= nimg) return;
    const float2 v = in[n + nimg * f];
    for (int c = 0; c < C; c++) {
        const float2 s = csm[n + nimg * c];
        out[n + nimg * (f + (long long)F * c)] = make_float2(v.x * s.x - v.y * s.y, v.x * s.y + v.y * s.x);
    }
}

// out[n, f] (+)= sum_c conj(csm[n, c]) * in[n, f, c]
__global__ void csm_conj_sum_kernel(const float2* __restrict__ in, float2* __restrict__ out,
                                    const float2* __restrict__ csm, long long nimg, int F, int C, int accumulate) {
    const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const int f = blockIdx.y;
    if (n >= nimg) return;
    float2 a = make_float2(0.0f, 0.0f);
    for (int c = 0; c < C; c++) {
        const float2 s = csm[n + nimg * c];
        const float2 v = in[n + nimg * (f + (long long)F * c)];
        a.x += s.x * v.x + s.y * v.y;
        a.y += s.x * v.y - s.y * v.x;
    }
    float2* o = out + n + nimg * f;
    if (accumulate) {
        a.x += o->x;
        a.y += o->y;
    }
    *o = a;
}

__global__ void add_kernel(float2* __restrict__ out, const float2* __restrict__ in, long long n) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = make_float2(out[i].x + in[i].x, out[i].y + in[i].y);
}

// ---------------------------------------------------------------------------
// CG state and FP64 reductions
// ---------------------------------------------------------------------------
struct CgState {
    double rq0, rq;
    double alpha_re, alpha_im;
    double beta;
    int done;  // set after the iteration where rq / rq0 < tolerance
    int it;
};

struct CgParams {
    long long n;            // image elements
    float we_re;            // encoding operator weight
    float kappa;            // identity regularisation weight
    const float* regd;      // or null: per element, kappa + sum_k weight_k image_k^2
    const float2* w;        // preconditioner weights or null
    long long wlen;
    double tol;
};

// Regulariser diagonal of the CG operator: out = (first ? base : out) +
// weight * image^2 (image broadcast by modulus; null: the identity).
__global__ void reg_diag_kernel(float* __restrict__ out, long long n, float base, const float* __restrict__ img,
                                long long len, float weight, int first) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        float w = 1.0f;
        if (img) w = img[i % len];
        const float t = weight * (w * w);  // imageOperator::mult_MH_M: x * image * image, then the weight
        out[i] = (first ? base : out[i]) + t;
    }
}

// Block reduction of a double2 in a fixed order.
__device__ __forceinline__ double2 block_sum(double2 v, double2* sh) {
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int s = kRedThreads / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            sh[threadIdx.x].x += sh[threadIdx.x + s].x;
            sh[threadIdx.x].y += sh[threadIdx.x + s].y;
        }
        __syncthreads();
    }
    return sh[0];
}

__device__ __forceinline__ float2 precond2(const CgParams& P, long long i, float2 v) {
    // cgPreconditioner::apply twice: v * w * w (weights broadcast by modulus)
    const float2 w = P.w[i % P.wlen];
    const float2 t = make_float2(v.x * w.x - v.y * w.y, v.x * w.y + v.y * w.x);
    return make_float2(t.x * w.x - t.y * w.y, t.x * w.y + t.y * w.x);
}

// Initialisation: r = rhs_scale * mhd; p = r (or P^2 r); x = 0 (x_mode 0),
// x = r (1: what a solve of zero iterations returns) or x untouched (2: an
// initial guess follows, which may be x itself); partial sums of <r, p>.
__device__ __forceinline__ void cg_init_run(const float2* __restrict__ mhd, float2* __restrict__ x,
                                                              float2* __restrict__ r, float2* __restrict__ p,
                                                              CgParams P, float rhs_scale, int x_mode,
                                                              double2* __restrict__ partial) {
    __shared__ double2 sh[kRedThreads];
    double2 acc = make_double2(0.0, 0.0);
    for (long long i = (long long)blockIdx.x * kRedThreads + threadIdx.x; i < P.n; i += (long long)gridDim.x * kRedThreads) {
        const float2 d = mhd[i];
        const float2 rv = make_float2(rhs_scale * d.x, rhs_scale * d.y);
        const float2 pv = P.w ? precond2(P, i, rv) : rv;
        r[i] = rv;
        p[i] = pv;
        if (x_mode != 2) x[i] = x_mode ? rv : make_float2(0.0f, 0.0f);
        acc.x += (double)rv.x * pv.x + (double)rv.y * pv.y;  // Re(conj(r) p)
    }
    const double2 s = block_sum(acc, sh);
    if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

__global__ __launch_bounds__(kRedThreads) void cg_init_kernel(const float2* __restrict__ mhd, float2* __restrict__ x,
                                                              float2* __restrict__ r, float2* __restrict__ p,
                                                              CgParams P, float rhs_scale, int x_mode,
                                                              double2* __restrict__ partial) {
    cg_init_run(mhd, x, r, p, P, rhs_scale, x_mode, partial);
}

__device__ __forceinline__ void cg_init_finish_run(const double2* __restrict__ partial, int nb,
                                                                     CgState* st, double* hist) {
    __shared__ double2 sh[kRedThreads];
    double2 acc = make_double2(0.0, 0.0);
    for (int i = threadIdx.x; i < nb; i += kRedThreads) {
        acc.x += partial[i].x;
        acc.y += partial[i].y;
    }
    const double2 s = block_sum(acc, sh);
    if (threadIdx.x == 0) {
        st->rq0 = s.x;
        st->rq = s.x;
        st->alpha_re = st->alpha_im = 0.0;
        st->beta = 0.0;
        st->done = !(s.x > 0.0);  // a zero right-hand side has the solution x = 0
        st->it = 0;
        if (hist) hist[0] = s.x;
    }
}

__global__ __launch_bounds__(kRedThreads) void cg_init_finish_kernel(const double2* __restrict__ partial, int nb,
                                                                     CgState* st, double* hist) {
    cg_init_finish_run(partial, nb, st, hist);
}

// q = w_E sum_c conj(csm_c) img_c + kappa p: the coil sum of E^H E p (in the
// order of csm_conj_sum_kernel) fused with the CG step; partial sums of <p, q>.
__device__ __forceinline__ void cg_q_csm_run(const float2* __restrict__ img,
                                                               const float2* __restrict__ csm, long long nimg, int F,
                                                               int C, const float2* __restrict__ p,
                                                               float2* __restrict__ q, CgParams P,
                                                               double2* __restrict__ partial) {
    __shared__ double2 sh[kRedThreads];
    double2 acc = make_double2(0.0, 0.0);
    for (long long i = (long long)blockIdx.x * kRedThreads + threadIdx.x; i < P.n; i += (long long)gridDim.x * kRedThreads) {
        const long long f = i / nimg, n = i - f * nimg;
        float2 a = make_float2(0.0f, 0.0f);
        for (int c = 0; c < C; c++) {
            const float2 s = csm[n + nimg * c];
            const float2 v = img[n + nimg * (f + (long long)F * c)];
            a.x += s.x * v.x + s.y * v.y;
            a.y += s.x * v.y - s.y * v.x;
        }
        const float2 pv = p[i];
        const float kd = P.regd ? P.regd[i] : P.kappa;
        const float2 qv = make_float2(P.we_re * a.x + kd * pv.x, P.we_re * a.y + kd * pv.y);
        q[i] = qv;
        acc.x += (double)pv.x * qv.x + (double)pv.y * qv.y;   // Re(conj(p) q)
        acc.y += (double)pv.x * qv.y - (double)pv.y * qv.x;   // Im(conj(p) q)
    }
    const double2 s = block_sum(acc, sh);
    if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

__global__ __launch_bounds__(kRedThreads) void cg_q_csm_kernel(const float2* __restrict__ img,
                                                               const float2* __restrict__ csm, long long nimg, int F,
                                                               int C, const float2* __restrict__ p,
                                                               float2* __restrict__ q, CgParams P,
                                                               double2* __restrict__ partial) {
    cg_q_csm_run(img, csm, nimg, F, C, p, q, P, partial);
}

// Initial guess (cgSolver::initialize with x0, cgSolver.h:254-277): x = x0,
// r -= A x0 (A x0 = w_E sum_c conj(csm_c) img_c + regd x0, img = E^H E x0
// before the coil sum), p = P^2 r; partial sums of <r, p>.
__global__ __launch_bounds__(kRedThreads) void cg_x0_kernel(const float2* __restrict__ img,
                                                            const float2* __restrict__ csm, long long nimg, int F,
                                                            int C, const float2* __restrict__ x0,
                                                            float2* __restrict__ x, float2* __restrict__ r,
                                                            float2* __restrict__ p, CgParams P,
                                                            double2* __restrict__ partial) {
    __shared__ double2 sh[kRedThreads];
    double2 acc = make_double2(0.0, 0.0);
    for (long long i = (long long)blockIdx.x * kRedThreads + threadIdx.x; i < P.n; i += (long long)gridDim.x * kRedThreads) {
        const long long f = i / nimg, n = i - f * nimg;
        float2 a = make_float2(0.0f, 0.0f);
        for (int c = 0; c < C; c++) {
            const float2 s = csm[n + nimg * c];
            const float2 v = img[n + nimg * (f + (long long)F * c)];
            a.x += s.x * v.x + s.y * v.y;
            a.y += s.x * v.y - s.y * v.x;
        }
        const float2 xv = x0[i];
        const float kd = P.regd ? P.regd[i] : P.kappa;
        float2 rv = r[i];
        rv.x -= P.we_re * a.x + kd * xv.x;
        rv.y -= P.we_re * a.y + kd * xv.y;
        const float2 pv = P.w ? precond2(P, i, rv) : rv;
        r[i] = rv;
        p[i] = pv;
        x[i] = xv;
        acc.x += (double)rv.x * pv.x + (double)rv.y * pv.y;
    }
    const double2 s = block_sum(acc, sh);
    if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// rq = Re <r, p> after the initial guess; rq_0 keeps the right-hand side's value.
__global__ __launch_bounds__(kRedThreads) void cg_x0_finish_kernel(const double2* __restrict__ partial, int nb,
                                                                   CgState* st) {
    __shared__ double2 sh[kRedThreads];
    double2 acc = make_double2(0.0, 0.0);
    for (int i = threadIdx.x; i < nb; i += kRedThreads) {
        acc.x += partial[i].x;
        acc.y += partial[i].y;
    }
    const double2 s = block_sum(acc, sh);
    if (threadIdx.x == 0) {
        st->rq = s.x;
        st->done = !(s.x > 0.0);  // x0 already solves the system
    }
}

// alpha = rq / <p, q> (complex); zero once the solver has terminated.
__global__ __launch_bounds__(kRedThreads) void cg_alpha_kernel(const double2* __restrict__ partial, int nb, CgState* st) {
    __shared__ double2 sh[kRedThreads];
    double2 acc = make_double2(0.0, 0.0);
    for (int i = threadIdx.x; i < nb; i += kRedThreads) {
        acc.x += partial[i].x;
        acc.y += partial[i].y;
    }
    const double2 s = block_sum(acc, sh);
    if (threadIdx.x == 0) {
        const double m = s.x * s.x + s.y * s.y;
        if (st->done || !(m > 0.0)) {
            st->alpha_re = st->alpha_im = 0.0;
        } else {
            st->alpha_re = st->rq * s.x / m;
            st->alpha_im = -st->rq * s.y / m;
        }
    }
}

// x += alpha p; r -= alpha q; then z = P^2 r into q (or z = r); partial sums of Re<r, z>.
__global__ __launch_bounds__(kRedThreads) void cg_xr_kernel(float2* __restrict__ x, float2* __restrict__ r,
                                                            const float2* __restrict__ p, float2* __restrict__ q,
                                                            CgParams P, const CgState* __restrict__ st,
                                                            double2* __restrict__ partial) {
    __shared__ double2 sh[kRedThreads];
    const float ar = (float)st->alpha_re, ai = (float)st->alpha_im;
    double2 acc = make_double2(0.0, 0.0);
    for (long long i = (long long)blockIdx.x * kRedThreads + threadIdx.x; i < P.n; i += (long long)gridDim.x * kRedThreads) {
        const float2 pv = p[i], qv = q[i];
        float2 xv = x[i], rv = r[i];
        xv.x += ar * pv.x - ai * pv.y;
        xv.y += ar * pv.y + ai * pv.x;
        rv.x -= ar * qv.x - ai * qv.y;
        rv.y -= ar * qv.y + ai * qv.x;
        x[i] = xv;
        r[i] = rv;
        float2 z = rv;
        if (P.w) {
            z = precond2(P, i, rv);
            q[i] = z;
        }
        acc.x += (double)rv.x * z.x + (double)rv.y * z.y;
    }
    const double2 s = block_sum(acc, sh);
    if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// beta = rq' / rq; rq = rq'; termination test (relativeResidualTCB).
__global__ __launch_bounds__(kRedThreads) void cg_beta_kernel(const double2* __restrict__ partial, int nb, CgState* st,
                                                              double* hist, double tol) {
    __shared__ double2 sh[kRedThreads];
    double2 acc = make_double2(0.0, 0.0);
    for (int i = threadIdx.x; i < nb; i += kRedThreads) {
        acc.x += partial[i].x;
        acc.y += partial[i].y;
    }
    const double2 s = block_sum(acc, sh);
    if (threadIdx.x == 0) {
        const int it = st->it;
        if (st->done) {
            st->beta = 0.0;
            if (hist) hist[it + 1] = 0.0;
        } else {
            st->beta = st->rq > 0.0 ? s.x / st->rq : 0.0;
            st->rq = s.x;
            if (hist) hist[it + 1] = s.x;
            if (s.x / st->rq0 < tol || !(s.x > 0.0)) st->done = 1;
        }
        st->it = it + 1;
    }
}

// p = beta p + z (z = P^2 r, held in q, or r)
// Device-side iteration it of the solve (no host reads): the state of
// iteration it is st[it], and the step kernels below each sum the FP64 partial
// sums of the previous kernel themselves, in the same order in every block, so
// the alpha and beta finishers are not separate launches.  x, r update with
// alpha from the <p, q> partials; partial sums of <r, z>.
__device__ __forceinline__ double2 sum_partials(const double2* __restrict__ partial, int nb, double2* sh) {
    double2 acc = make_double2(0.0, 0.0);
    for (int i = threadIdx.x; i < nb; i += kRedThreads) {
        acc.x += partial[i].x;
        acc.y += partial[i].y;
    }
    const double2 s = block_sum(acc, sh);
    __syncthreads();  // sh is reused by the caller
    return s;
}

__device__ __forceinline__ void cg_xr_step_run(float2* __restrict__ x, float2* __restrict__ r,
                                                                 const float2* __restrict__ p, float2* __restrict__ q,
                                                                 CgParams P, const CgState* __restrict__ st,
                                                                 const double2* __restrict__ pq, int nb,
                                                                 double2* __restrict__ partial) {
    __shared__ double2 sh[kRedThreads];
    const double2 s = sum_partials(pq, nb, sh);
    const double m = s.x * s.x + s.y * s.y;
    float ar = 0.0f, ai = 0.0f;
    if (!st->done && m > 0.0) {  // cg_alpha_kernel
        ar = (float)(st->rq * s.x / m);
        ai = (float)(-st->rq * s.y / m);
    }
    double2 acc = make_double2(0.0, 0.0);
    for (long long i = (long long)blockIdx.x * kRedThreads + threadIdx.x; i < P.n; i += (long long)gridDim.x * kRedThreads) {
        const float2 pv = p[i], qv = q[i];
        float2 xv = x[i], rv = r[i];
        xv.x += ar * pv.x - ai * pv.y;
        xv.y += ar * pv.y + ai * pv.x;
        rv.x -= ar * qv.x - ai * qv.y;
        rv.y -= ar * qv.y + ai * qv.x;
        x[i] = xv;
        r[i] = rv;
        float2 z = rv;
        if (P.w) {
            z = precond2(P, i, rv);
            q[i] = z;
        }
        acc.x += (double)rv.x * z.x + (double)rv.y * z.y;
    }
    const double2 t = block_sum(acc, sh);
    if (threadIdx.x == 0) partial[blockIdx.x] = t;
}

__global__ __launch_bounds__(kRedThreads) void cg_xr_step_kernel(float2* __restrict__ x, float2* __restrict__ r,
                                                                 const float2* __restrict__ p, float2* __restrict__ q,
                                                                 CgParams P, const CgState* __restrict__ st,
                                                                 const double2* __restrict__ pq, int nb,
                                                                 double2* __restrict__ partial) {
    cg_xr_step_run(x, r, p, q, P, st, pq, nb, partial);
}

// p = z + beta p with beta from the <r, z> partials; block 0 writes the state
// of the next iteration (cg_beta_kernel).
__device__ __forceinline__ void cg_p_step_run(float2* __restrict__ p, const float2* __restrict__ z,
                                                                long long n, const CgState* __restrict__ st,
                                                                CgState* __restrict__ next,
                                                                const double2* __restrict__ rz, int nb,
                                                                double* __restrict__ hist, int it, double tol) {
    __shared__ double2 sh[kRedThreads];
    const double rr = sum_partials(rz, nb, sh).x;
    const CgState c = *st;
    const double beta = c.done ? 0.0 : (c.rq > 0.0 ? rr / c.rq : 0.0);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        CgState nx = c;
        nx.alpha_re = nx.alpha_im = 0.0;
        nx.beta = beta;
        nx.it = it + 1;
        if (!c.done) {
            nx.rq = rr;
            if (rr / c.rq0 < tol || !(rr > 0.0)) nx.done = 1;
        }
        *next = nx;
        if (hist) hist[it + 1] = c.done ? 0.0 : rr;
    }
    const float b = (float)beta;
    for (long long i = (long long)blockIdx.x * kRedThreads + threadIdx.x; i < n; i += (long long)gridDim.x * kRedThreads) {
        const float2 pv = p[i], zv = z[i];
        p[i] = make_float2(b * pv.x + zv.x, b * pv.y + zv.y);
    }
}

__global__ __launch_bounds__(kRedThreads) void cg_p_step_kernel(float2* __restrict__ p, const float2* __restrict__ z,
                                                                long long n, const CgState* __restrict__ st,
                                                                CgState* __restrict__ next,
                                                                const double2* __restrict__ rz, int nb,
                                                                double* __restrict__ hist, int it, double tol) {
    cg_p_step_run(p, z, n, st, next, rz, nb, hist, it, tol);
}

__global__ void cg_p_kernel(float2* __restrict__ p, const float2* __restrict__ z, long long n,
                            const CgState* __restrict__ st) {
    const float b = (float)st->beta;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        const float2 pv = p[i], zv = z[i];
        p[i] = make_float2(b * pv.x + zv.x, b * pv.y + zv.y);
    }
}

// ---------------------------------------------------------------------------
// The fused batch: one launch per step for every problem, grid (blocks,
// problems), blockIdx.y selecting the problem.  Each problem's step runs the
// same body as the single-problem kernel over its own vectors, partial sums
// and state.
// ---------------------------------------------------------------------------
constexpr int kFusedRedBlocks = 64;  // partial sums per reduction and problem

struct CgProb {
    float2* x;
    float2* r;
    float2* p;
    float2* q;
    float2* mhd;
    const float2* coil;  // coil images of E^H (before the coil sum), [N..., F, C]
    const float2* csm;
    long long nimg;
    int F, C;
    CgParams P;
    CgState* st;    // iterations + 1 states
    double2* part;  // 2 kFusedRedBlocks partial sums
};

// mhd = sum_c conj(csm_c) coil_c (the coil sum of E^H d), as csm_conj_sum_kernel.
__global__ __launch_bounds__(kRedThreads) void cg_coil_sum_batch_kernel(const CgProb* __restrict__ probs) {
    const CgProb d = probs[blockIdx.y];
    for (long long i = (long long)blockIdx.x * kRedThreads + threadIdx.x; i < d.P.n;
         i += (long long)gridDim.x * kRedThreads) {
        const long long f = i / d.nimg, n = i - f * d.nimg;
        float2 a = make_float2(0.0f, 0.0f);
        for (int c = 0; c < d.C; c++) {
            const float2 s = d.csm[n + d.nimg * c];
            const float2 v = d.coil[n + d.nimg * (f + (long long)d.F * c)];
            a.x += s.x * v.x + s.y * v.y;
            a.y += s.x * v.y - s.y * v.x;
        }
        d.mhd[i] = a;
    }
}

__global__ __launch_bounds__(kRedThreads) void cg_init_batch_kernel(const CgProb* __restrict__ probs) {
    const CgProb d = probs[blockIdx.y];
    cg_init_run(d.mhd, d.x, d.r, d.p, d.P, d.P.we_re, 0, d.part);
}

__global__ __launch_bounds__(kRedThreads) void cg_init_finish_batch_kernel(const CgProb* __restrict__ probs) {
    const CgProb d = probs[blockIdx.y];
    cg_init_finish_run(d.part, kFusedRedBlocks, d.st, nullptr);
}

__global__ __launch_bounds__(kRedThreads) void cg_q_csm_batch_kernel(const CgProb* __restrict__ probs) {
    const CgProb d = probs[blockIdx.y];
    cg_q_csm_run(d.coil, d.csm, d.nimg, d.F, d.C, d.p, d.q, d.P, d.part);
}

__global__ __launch_bounds__(kRedThreads) void cg_xr_step_batch_kernel(const CgProb* __restrict__ probs, int it) {
    const CgProb d = probs[blockIdx.y];
    cg_xr_step_run(d.x, d.r, d.p, d.q, d.P, d.st + it, d.part, kFusedRedBlocks, d.part + kFusedRedBlocks);
}

__global__ __launch_bounds__(kRedThreads) void cg_p_step_batch_kernel(const CgProb* __restrict__ probs, int it) {
    const CgProb d = probs[blockIdx.y];
    cg_p_step_run(d.p, d.r, d.P.n, d.st + it, d.st + it + 1, d.part + kFusedRedBlocks, kFusedRedBlocks, nullptr, it,
                  d.P.tol);
}

// ---------------------------------------------------------------------------
// imageOperator / cuImageOperator::compute (operators/imageOperator.h:28-47,
// operators/gpu/cuImageOperator.h:30-63) and the CG-SENSE gadget's
// preconditioner weights (gadgets/pmri/gpuCgSenseGadget.cpp:165-176)
// ---------------------------------------------------------------------------
constexpr int kHistBins = 50000;  // cuImageOperator::estimate_offset granularity

// FP64 partial sums of |re| + |im| (asum, the BLAS definition for complex data).
__global__ __launch_bounds__(kRedThreads) void asum_kernel(const float2* __restrict__ x, long long n,
                                                           double2* __restrict__ partial) {
    __shared__ double2 sh[kRedThreads];
    double2 acc = make_double2(0.0, 0.0);
    for (long long i = (long long)blockIdx.x * kRedThreads + threadIdx.x; i < n; i += (long long)gridDim.x * kRedThreads) {
        const float2 v = x[i];
        acc.x += (double)fabsf(v.x) + (double)fabsf(v.y);
    }
    const double2 s = block_sum(acc, sh);
    if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// scal[0] = n / asum (float, as the reference's REAL arithmetic), scal[2] = asum.
__global__ __launch_bounds__(kRedThreads) void asum_finish_kernel(const double2* __restrict__ partial, int nb,
                                                                  long long n, float* __restrict__ scal) {
    __shared__ double2 sh[kRedThreads];
    double2 acc = make_double2(0.0, 0.0);
    for (int i = threadIdx.x; i < nb; i += kRedThreads) {
        acc.x += partial[i].x;
        acc.y += partial[i].y;
    }
    const double2 s = block_sum(acc, sh);
    if (threadIdx.x == 0) {
        const float sum = (float)s.x;
        scal[0] = (float)n / sum;
        scal[2] = sum;
    }
}

// out = |x * scale| (tmp *= scale; image_ = abs(tmp)); per-block maxima.
__global__ __launch_bounds__(kRedThreads) void abs_scale_kernel(const float2* __restrict__ x, long long n,
                                                                const float* __restrict__ scal, float* __restrict__ out,
                                                                float* __restrict__ bmax) {
    __shared__ float sh[kRedThreads];
    const float scale = scal[0];
    float m = 0.0f;
    for (long long i = (long long)blockIdx.x * kRedThreads + threadIdx.x; i < n; i += (long long)gridDim.x * kRedThreads) {
        const float2 v = x[i];
        const float re = v.x * scale, im = v.y * scale;
        const float a = sqrtf(re * re + im * im);
        out[i] = a;
        m = fmaxf(m, a);
    }
    sh[threadIdx.x] = m;
    __syncthreads();
    for (int t = kRedThreads / 2; t > 0; t >>= 1) {
        if ((int)threadIdx.x < t) sh[threadIdx.x] = fmaxf(sh[threadIdx.x], sh[threadIdx.x + t]);
        __syncthreads();
    }
    if (threadIdx.x == 0) bmax[blockIdx.x] = sh[0];
}

// scal[1] = the image maximum (amax of a non-negative image).
__global__ __launch_bounds__(kRedThreads) void max_finish_kernel(const float* __restrict__ bmax, int nb,
                                                                 float* __restrict__ scal) {
    __shared__ float sh[kRedThreads];
    float m = 0.0f;
    for (int i = threadIdx.x; i < nb; i += kRedThreads) m = fmaxf(m, bmax[i]);
    sh[threadIdx.x] = m;
    __syncthreads();
    for (int t = kRedThreads / 2; t > 0; t >>= 1) {
        if ((int)threadIdx.x < t) sh[threadIdx.x] = fmaxf(sh[threadIdx.x], sh[threadIdx.x + t]);
        __syncthreads();
    }
    if (threadIdx.x == 0) scal[1] = sh[0];
}

__global__ void zero_u32_kernel(unsigned int* __restrict__ y, long long n) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        y[i] = 0u;
}

// bin = min(floor(value / max * 50000), 49999); integer counts (order free).
__global__ void hist_kernel(const float* __restrict__ img, long long n, const float* __restrict__ scal,
                            unsigned int* __restrict__ bins) {
    const float mx = scal[1];
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        const float b = floorf((img[i] / mx) * (float)kHistBins);
        const unsigned int k = min((unsigned int)b, (unsigned int)(kHistBins - 1));
        atomicAdd(&bins[k], 1u);
    }
}

// image_ += offset (when positive); reciprocal_inplace.
__global__ void offset_recip_kernel(float* __restrict__ img, long long n, float offset) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        float v = img[i];
        if (offset > 0.0f) v += offset;
        img[i] = 1.0f / v;
    }
}

// w = 1 / sqrt(sum_c |csm_c|^2 + kappa reg), as complex (w, 0).
__global__ void precond_weights_kernel(const float2* __restrict__ csm, long long nimg, int C,
                                       const float* __restrict__ reg, float kappa, float2* __restrict__ w) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < nimg; i += (long long)gridDim.x * blockDim.x) {
        float a = 0.0f;
        for (int c = 0; c < C; c++) {
            const float2 v = csm[i + nimg * c];
            a += v.x * v.x + v.y * v.y;
        }
        if (reg) a += kappa * reg[i];
        w[i] = make_float2(1.0f / sqrtf(a), 0.0f);
    }
}

unsigned blocks_for(long long n, int threads) {
    return (unsigned)std::max<long long>(1, std::min<long long>((n + threads - 1) / threads, 1 << 20));
}

}  // namespace

// ===========================================================================
// operator and solver objects
// ===========================================================================
struct gt_sense_op {
    gt_nufft_plan* plan = nullptr;
    int D = 0;
    size_t N[3] = {1, 1, 1};
    size_t nimg = 0;  // prod N
    size_t M = 0, F = 0, C = 0;
    int device = 0;
    const float2* csm = nullptr;
    const float* dcw = nullptr;
    size_t dcw_len = 0;
    DevBuf coil_img;   // [N..., F, C]
    DevBuf coil_smp;   // [M, F, C]
    // CG workspace
    DevBuf r, p, q, mhd;
    DevBuf regd;  // regulariser diagonal (image regularisers)
    DevBuf partial, state, hist;
    bool warmed = false;
    std::pair<uint64_t, uint64_t> warm_gen{0, 0};  // (plan, operator) generations at the last warm
    // bumped by every reallocation of the buffers above (their epoch points
    // here): a recorded batch graph holds their addresses
    uint64_t generation = 0;
    void track() {
        for (DevBuf* b : {&coil_img, &coil_smp, &r, &p, &q, &mhd, &regd, &partial, &state, &hist}) b->epoch = &generation;
    }

    size_t image_elems() const { return nimg * F; }

    void img_dims(size_t* d, int& nd, bool coils) const {
        nd = 0;
        for (int i = 0; i < D; i++) d[nd++] = N[i];
        d[nd++] = F;
        if (coils) d[nd++] = C;
    }

    void check(gt_status s) const {
        if (s != GT_OK) throw gt_error(s, last_error());
    }

    void mult_M(const void* in, void* out, bool accumulate, hipStream_t s) {
        void* dst = accumulate ? coil_smp.ptr : out;
        // the plan multiplies the coil maps inside its first FFT pass where it
        // has one; otherwise the coil images go through coil_img
        if (!forward_with_coil_maps(plan, in, csm, C, dst, dcw, dcw_len, s)) {
            size_t di[5], ds[3] = {M, F, C};
            int ni;
            img_dims(di, ni, true);
            dim3 grid(blocks_for((long long)nimg, 256), (unsigned)F);
            hipLaunchKernelGGL(csm_mult_kernel, grid, dim3(256), 0, s, (const float2*)in, coil_img.as<float2>(), csm,
                               (long long)nimg, int(F), int(C));
            GT_HIP(hipGetLastError());
            check(gt_nufft_compute(plan, coil_img.ptr, di, ni, dst, ds, 3, dcw, dcw_len, GT_FORWARDS_C2NC, s));
        }
        if (accumulate) {
            const long long n = (long long)(M * F * C);
            hipLaunchKernelGGL(add_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, s, (float2*)out,
                               coil_smp.as<float2>(), n);
            GT_HIP(hipGetLastError());
        }
    }

    void mult_MH(const void* in, void* out, bool accumulate, hipStream_t s) {
        size_t di[5], ds[3] = {M, F, C};
        int ni;
        img_dims(di, ni, true);
        check(gt_nufft_compute(plan, in, ds, 3, coil_img.ptr, di, ni, dcw, dcw_len, GT_BACKWARDS_NC2C, s));
        dim3 grid(blocks_for((long long)nimg, 256), (unsigned)F);
        hipLaunchKernelGGL(csm_conj_sum_kernel, grid, dim3(256), 0, s, coil_img.as<float2>(), (float2*)out, csm,
                           (long long)nimg, int(F), int(C), accumulate ? 1 : 0);
        GT_HIP(hipGetLastError());
    }

    // E^H E in, leaving the coil images E^H E before the coil sum in coil_img
    void mult_MH_M_coils(const void* in, hipStream_t s) {
        mult_M(in, coil_smp.ptr, false, s);
        // coil_smp is consumed by NC2C before mult_MH writes coil_img
        size_t di[5], ds[3] = {M, F, C};
        int ni;
        img_dims(di, ni, true);
        check(gt_nufft_compute(plan, coil_smp.ptr, ds, 3, coil_img.ptr, di, ni, dcw, dcw_len, GT_BACKWARDS_NC2C, s));
    }

    void mult_MH_M(const void* in, void* out, bool accumulate, hipStream_t s) {
        mult_MH_M_coils(in, s);
        dim3 grid(blocks_for((long long)nimg, 256), (unsigned)F);
        hipLaunchKernelGGL(csm_conj_sum_kernel, grid, dim3(256), 0, s, coil_img.as<float2>(), (float2*)out, csm,
                           (long long)nimg, int(F), int(C), accumulate ? 1 : 0);
        GT_HIP(hipGetLastError());
    }

    void reserve_cg(int iterations, bool regs = false) {
        const size_t bytes = image_elems() * sizeof(float2);
        r.reserve(bytes);
        p.reserve(bytes);
        q.reserve(bytes);
        mhd.reserve(bytes);
        if (regs) regd.reserve(image_elems() * sizeof(float));
        partial.reserve(2 * kRedBlocks * sizeof(double2));
        state.reserve(size_t(iterations + 1) * sizeof(CgState));  // one per iteration (device-side solve)
        hist.reserve(size_t(iterations + 1) * sizeof(double));
    }

    // One E^H E application grows every buffer the solve uses (the plan's
    // grid, partial tiles, ...), so a later graph capture allocates nothing.
    // It runs again whenever the plan or these buffers changed since.
    void warm(hipStream_t s) {
        const auto now = std::make_pair(plan_generation(plan), generation);
        if (warmed && now == warm_gen) return;
        mult_MH_M(p.ptr, q.ptr, false, s);
        GT_HIP(hipStreamSynchronize(s));
        warmed = true;
        warm_gen = std::make_pair(plan_generation(plan), generation);
    }
};

namespace {

CgParams cg_params(const gt_sense_op* op, const gt_cg_desc* d) {
    CgParams P{};
    P.n = (long long)op->image_elems();
    P.we_re = float(d->encoding_weight);
    P.kappa = float(d->kappa);
    P.regd = d->n_regs > 0 ? op->regd.as<float>() : nullptr;
    P.w = static_cast<const float2*>(d->precond);
    P.wlen = (long long)d->precond_len;
    P.tol = d->tc_tolerance;
    return P;
}

void validate_desc(const gt_cg_desc* d) {
    GT_REQUIRE(d, GT_ERR_INVALID_ARGUMENT, "null CG descriptor");
    GT_REQUIRE(d->iterations >= 0, GT_ERR_INVALID_ARGUMENT, "negative iteration count");
    GT_REQUIRE(!d->precond || d->precond_len > 0, GT_ERR_INVALID_ARGUMENT, "preconditioner with zero length");
    GT_REQUIRE(d->n_regs >= 0 && (d->n_regs == 0 || d->regs), GT_ERR_INVALID_ARGUMENT,
               "regularisers: negative count or null list");
    for (int k = 0; k < d->n_regs; k++)
        GT_REQUIRE(!d->regs[k].image || d->regs[k].image_len > 0, GT_ERR_INVALID_ARGUMENT,
                   "regulariser image with zero length");
}

// Sizes of the descriptor's arrays against the operator's image.
void validate_sizes(const gt_sense_op* op, const gt_cg_desc* d) {
    const long long n = (long long)op->image_elems();
    GT_REQUIRE(!d->precond || n % (long long)d->precond_len == 0, GT_ERR_SHAPE,
               "preconditioner length does not divide the image size");
    for (int k = 0; k < d->n_regs; k++)
        GT_REQUIRE(!d->regs[k].image || n % (long long)d->regs[k].image_len == 0, GT_ERR_SHAPE,
                   "regulariser image length does not divide the image size");
}

// regd = kappa + sum_k weight_k image_k^2, in list order (cgSolver::mult_MH_M
// adds the regularisers after the encoding operator, cgSolver.h:376-384).
void build_regd(gt_sense_op* op, const gt_cg_desc* d, hipStream_t s) {
    if (d->n_regs <= 0) return;
    const long long n = (long long)op->image_elems();
    for (int k = 0; k < d->n_regs; k++) {
        const gt_cg_reg& g = d->regs[k];
        hipLaunchKernelGGL(reg_diag_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, s, op->regd.as<float>(), n,
                           float(d->kappa), g.image, (long long)std::max<size_t>(g.image_len, 1), float(g.weight),
                           k == 0 ? 1 : 0);
    }
    GT_HIP(hipGetLastError());
}

// Host callback of the plain solver: the reference reads every dot product
// back (cgSolver.h:333-365 through cuNDArray_blas dot); here each finished
// reduction is copied to the host and waited for.
void host_read(const gt_sense_op* op, CgState* h, hipStream_t s) {
    GT_HIP(hipMemcpyAsync(h, op->state.ptr, sizeof(CgState), hipMemcpyDeviceToHost, s));
    GT_HIP(hipStreamSynchronize(s));
}

// The whole solve on one stream.  host_sync: read the scalars back after each
// reduction (plain method); otherwise no host round trip (capturable).
// image_rhs: `data` is the right-hand side itself, an image [N..., F], instead
// of samples whose w_E E^H image is the right-hand side.
void cg_solve(gt_sense_op* op, const void* data, void* x, const gt_cg_desc* d, bool host_sync, hipStream_t s,
              bool image_rhs = false) {
    validate_sizes(op, d);
    const CgParams P = cg_params(op, d);
    build_regd(op, d, s);
    const unsigned nb = (unsigned)std::min<long long>(kRedBlocks, (P.n + kRedThreads - 1) / kRedThreads);
    double2* part = op->partial.as<double2>();
    CgState* st = op->state.as<CgState>();
    double* hist = op->hist.as<double>();
    float2* xf = static_cast<float2*>(x);
    CgState h{};
    // rhs = w_E E^H d (cgSolver::compute_rhs), x = 0, r = p = rhs (initialize)
    if (image_rhs)
        GT_HIP(hipMemcpyAsync(op->mhd.ptr, data, op->image_elems() * sizeof(float2), hipMemcpyDeviceToDevice, s));
    else
        op->mult_MH(data, op->mhd.ptr, false, s);
    hipLaunchKernelGGL(cg_init_kernel, dim3(nb), dim3(kRedThreads), 0, s, op->mhd.as<float2>(), xf,
                       op->r.as<float2>(), op->p.as<float2>(), P, image_rhs ? 1.0f : P.we_re,
                       d->iterations == 0 ? 1 : (d->x0 ? 2 : 0), part);
    hipLaunchKernelGGL(cg_init_finish_kernel, dim3(1), dim3(kRedThreads), 0, s, part, int(nb), st, hist);
    GT_HIP(hipGetLastError());
    if (d->x0 && d->iterations > 0) {
        // x = x0, r = rhs - A x0, p = P^2 r (rq_0 stays the right-hand side's)
        op->mult_MH_M_coils(d->x0, s);
        hipLaunchKernelGGL(cg_x0_kernel, dim3(nb), dim3(kRedThreads), 0, s, op->coil_img.as<float2>(), op->csm,
                           (long long)op->nimg, int(op->F), int(op->C), static_cast<const float2*>(d->x0), xf,
                           op->r.as<float2>(), op->p.as<float2>(), P, part);
        hipLaunchKernelGGL(cg_x0_finish_kernel, dim3(1), dim3(kRedThreads), 0, s, part, int(nb), st);
        GT_HIP(hipGetLastError());
    }
    if (!host_sync) {
        // device-side: iteration it reads state it and writes state it + 1
        for (int it = 0; it < d->iterations; it++) {
            op->mult_MH_M_coils(op->p.ptr, s);
            hipLaunchKernelGGL(cg_q_csm_kernel, dim3(nb), dim3(kRedThreads), 0, s, op->coil_img.as<float2>(), op->csm,
                               (long long)op->nimg, int(op->F), int(op->C), op->p.as<float2>(), op->q.as<float2>(), P,
                               part);
            hipLaunchKernelGGL(cg_xr_step_kernel, dim3(nb), dim3(kRedThreads), 0, s, xf, op->r.as<float2>(),
                               op->p.as<float2>(), op->q.as<float2>(), P, st + it, part, int(nb), part + nb);
            const float2* z = P.w ? op->q.as<float2>() : op->r.as<float2>();
            hipLaunchKernelGGL(cg_p_step_kernel, dim3(nb), dim3(kRedThreads), 0, s, op->p.as<float2>(), z, P.n,
                               st + it, st + it + 1, part + nb, int(nb), hist, it, P.tol);
            GT_HIP(hipGetLastError());
        }
        return;
    }
    host_read(op, &h, s);
    if (h.done) return;
    for (int it = 0; it < d->iterations; it++) {
        // q = (w_E E^H E + kappa I) p
        op->mult_MH_M_coils(op->p.ptr, s);
        hipLaunchKernelGGL(cg_q_csm_kernel, dim3(nb), dim3(kRedThreads), 0, s, op->coil_img.as<float2>(), op->csm,
                           (long long)op->nimg, int(op->F), int(op->C), op->p.as<float2>(), op->q.as<float2>(), P,
                           part);
        hipLaunchKernelGGL(cg_alpha_kernel, dim3(1), dim3(kRedThreads), 0, s, part, int(nb), st);
        GT_HIP(hipGetLastError());
        if (host_sync) host_read(op, &h, s);
        hipLaunchKernelGGL(cg_xr_kernel, dim3(nb), dim3(kRedThreads), 0, s, xf, op->r.as<float2>(),
                           op->p.as<float2>(), op->q.as<float2>(), P, st, part);
        hipLaunchKernelGGL(cg_beta_kernel, dim3(1), dim3(kRedThreads), 0, s, part, int(nb), st, hist, P.tol);
        GT_HIP(hipGetLastError());
        if (host_sync) {
            host_read(op, &h, s);
            if (h.done) return;  // relativeResidualTCB: terminate
        }
        const float2* z = P.w ? op->q.as<float2>() : op->r.as<float2>();
        hipLaunchKernelGGL(cg_p_kernel, dim3(blocks_for(P.n, 256)), dim3(256), 0, s, op->p.as<float2>(), z, P.n, st);
        GT_HIP(hipGetLastError());
    }
}

}  // namespace

struct gt_cg_batch {
    std::vector<gt_sense_op*> ops;
    std::vector<const void*> data;
    std::vector<void*> x;
    gt_cg_desc desc{};
    std::vector<gt_cg_reg> regs;  // desc.regs points here (the caller's list may go away)
    int mode = GT_CG_PLAIN;
    int device = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int branches = 0;  // chosen by tune_graph
    // generations of each problem's plan and operator when the graph was
    // recorded: a replay after either changed would use stale addresses and
    // launch sizes
    std::vector<std::pair<uint64_t, uint64_t>> gens;
    std::vector<hipStream_t> streams;
    std::vector<hipEvent_t> events;
    // GT_CG_FUSED: the fused gridding of every problem, the CG vectors (r, p,
    // q, mhd of every problem in one allocation), states and partial sums,
    // and the per-problem table the batched CG kernels read
    FusedGridding* fused = nullptr;
    DevBuf fvec, fstate, fpart, fprobs;
    int fused_item_len = 0;  // NC2C entries per work item the autotuner chose
    ~gt_cg_batch() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        for (auto e : events) (void)hipEventDestroy(e);
        for (auto s : streams) (void)hipStreamDestroy(s);
        if (fused) fused_destroy(fused);
    }
};

namespace {

// Problems longest first onto the least loaded branch.  Cost of a solve ~
// samples x coils + grid x coils (bandwidth) plus a fixed term for its ~150
// launches (small problems are latency bound).
std::vector<std::vector<int>> assign_branches(const std::vector<gt_sense_op*>& ops, int nbr) {
    std::vector<int> order(ops.size());
    std::iota(order.begin(), order.end(), 0);
    auto work = [&](int i) {
        const gt_sense_op* o = ops[size_t(i)];
        return double(o->M * o->F * o->C) + 2.25 * double(o->nimg * o->F * o->C) + 4.0e5;
    };
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return work(a) > work(b); });
    std::vector<std::vector<int>> br(static_cast<size_t>(nbr));
    std::vector<double> load(static_cast<size_t>(nbr), 0.0);
    for (int i : order) {
        const size_t k = size_t(std::min_element(load.begin(), load.end()) - load.begin());
        br[k].push_back(i);
        load[k] += work(i);
    }
    return br;
}

// Captures one graph with nbr parallel branches, on a stream of its own (the
// caller's may be the null stream, which cannot be captured): fork from the
// capture stream, each branch runs its problems' whole solves back to back,
// join.  The instantiated graph is launched on the caller's stream.
void build_graph(gt_cg_batch* b, int nbr, hipGraph_t* graph, hipGraphExec_t* exec) {
    auto br = assign_branches(b->ops, nbr);
    while (int(b->streams.size()) < nbr + 1) {
        hipStream_t st;
        GT_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        b->streams.push_back(st);
    }
    while (int(b->events.size()) < nbr + 1) {
        hipEvent_t e;
        GT_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        b->events.push_back(e);
    }
    hipStream_t s = b->streams[size_t(nbr)];
    GT_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    GT_HIP(hipEventRecord(b->events[0], s));
    for (int k = 0; k < nbr; k++) {
        hipStream_t bs = b->streams[size_t(k)];
        GT_HIP(hipStreamWaitEvent(bs, b->events[0], 0));
        for (int i : br[size_t(k)]) cg_solve(b->ops[size_t(i)], b->data[size_t(i)], b->x[size_t(i)], &b->desc, false, bs);
        GT_HIP(hipEventRecord(b->events[size_t(k) + 1], bs));
        GT_HIP(hipStreamWaitEvent(s, b->events[size_t(k) + 1], 0));
    }
    GT_HIP(hipStreamEndCapture(s, graph));
    GT_HIP(hipGraphInstantiate(exec, *graph, nullptr, nullptr, 0));
}

// Autotuner: the number of parallel branches that runs the batch fastest.
// Each candidate graph is built and replayed twice (the second replay timed on
// the host around a synchronisation); the fastest is kept.
void tune_graph(gt_cg_batch* b, hipStream_t s) {
    const int n = int(b->ops.size());
    std::vector<int> cand;
    for (int c : {2, 4, 8, 16, 32, 64})
        if (c < n) cand.push_back(c);
    cand.push_back(n);  // one branch per problem
    if (cand.empty()) cand.push_back(1);
    double best = 1e300;
    for (int nbr : cand) {
        hipGraph_t g = nullptr;
        hipGraphExec_t e = nullptr;
        build_graph(b, nbr, &g, &e);
        GT_HIP(hipGraphLaunch(e, s));
        GT_HIP(hipStreamSynchronize(s));
        const auto t0 = std::chrono::steady_clock::now();
        GT_HIP(hipGraphLaunch(e, s));
        GT_HIP(hipGraphLaunch(e, s));
        GT_HIP(hipStreamSynchronize(s));
        const double t = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        if (t < best) {
            best = t;
            if (b->exec) (void)hipGraphExecDestroy(b->exec);
            if (b->graph) (void)hipGraphDestroy(b->graph);
            b->graph = g;
            b->exec = e;
            b->branches = nbr;
        } else {
            (void)hipGraphExecDestroy(e);
            (void)hipGraphDestroy(g);
        }
    }
}

// GT_CG_FUSED: the whole batch's solve, one launch per step for every problem
// (per size group for the FFT passes); capturable (no allocation, no host read).
void fused_solve(gt_cg_batch* b, hipStream_t s) {
    const unsigned np = (unsigned)b->ops.size();
    const CgProb* probs = b->fprobs.as<CgProb>();
    const dim3 grid(kFusedRedBlocks, np);
    // rhs = w_E E^H d, x = 0, r = p = rhs
    fused_adjoint(b->fused, true, s);
    hipLaunchKernelGGL(cg_coil_sum_batch_kernel, grid, dim3(kRedThreads), 0, s, probs);
    hipLaunchKernelGGL(cg_init_batch_kernel, grid, dim3(kRedThreads), 0, s, probs);
    hipLaunchKernelGGL(cg_init_finish_batch_kernel, dim3(1, np), dim3(kRedThreads), 0, s, probs);
    GT_HIP(hipGetLastError());
    for (int it = 0; it < b->desc.iterations; it++) {
        fused_forward(b->fused, s);         // E p, every problem
        fused_adjoint(b->fused, false, s);  // E^H E p before the coil sums
        hipLaunchKernelGGL(cg_q_csm_batch_kernel, grid, dim3(kRedThreads), 0, s, probs);
        hipLaunchKernelGGL(cg_xr_step_batch_kernel, grid, dim3(kRedThreads), 0, s, probs, it);
        hipLaunchKernelGGL(cg_p_step_batch_kernel, grid, dim3(kRedThreads), 0, s, probs, it);
        GT_HIP(hipGetLastError());
    }
}

// Builds the fused gridding and the CG tables of a GT_CG_FUSED batch, then
// records its solve into b->graph / b->exec.
void fused_setup(gt_cg_batch* b, hipStream_t s) {
    const int n = int(b->ops.size());
    std::vector<size_t> off(size_t(n) + 1, 0);
    for (int i = 0; i < n; i++) off[size_t(i) + 1] = off[size_t(i)] + b->ops[size_t(i)]->image_elems();
    const size_t total = off[size_t(n)];
    b->fvec.reserve(std::max<size_t>(4 * total, 1) * sizeof(float2));
    b->fstate.reserve(size_t(n) * size_t(b->desc.iterations + 1) * sizeof(CgState));
    b->fpart.reserve(size_t(n) * 2 * kFusedRedBlocks * sizeof(double2));
    float2* v = b->fvec.as<float2>();
    auto vec = [&](int i, int k) { return v + size_t(k) * total + off[size_t(i)]; };  // k: r, p, q, mhd
    std::vector<FusedProblem> fp;
    for (int i = 0; i < n; i++) {
        gt_sense_op* op = b->ops[size_t(i)];
        fp.push_back(FusedProblem{op->plan, op->C, op->dcw, op->dcw_len, op->csm, vec(i, 1), b->data[size_t(i)]});
    }
    if (b->fused) fused_destroy(b->fused);
    b->fused = nullptr;
    b->fused = fused_create(fp, s);
    std::vector<CgProb> probs(static_cast<size_t>(n));
    for (int i = 0; i < n; i++) {
        gt_sense_op* op = b->ops[size_t(i)];
        CgProb& d = probs[size_t(i)];
        d.x = static_cast<float2*>(b->x[size_t(i)]);
        d.r = vec(i, 0);
        d.p = vec(i, 1);
        d.q = vec(i, 2);
        d.mhd = vec(i, 3);
        d.coil = static_cast<const float2*>(fused_coil_images(b->fused, i));
        d.csm = op->csm;
        d.nimg = (long long)op->nimg;
        d.F = int(op->F);
        d.C = int(op->C);
        d.P = cg_params(op, &b->desc);
        d.st = b->fstate.as<CgState>() + size_t(i) * size_t(b->desc.iterations + 1);
        d.part = b->fpart.as<double2>() + size_t(i) * 2 * kFusedRedBlocks;
    }
    b->fprobs.reserve(size_t(n) * sizeof(CgProb));
    GT_HIP(hipMemcpyAsync(b->fprobs.ptr, probs.data(), size_t(n) * sizeof(CgProb), hipMemcpyHostToDevice, s));
    GT_HIP(hipStreamSynchronize(s));
    // record the solve on a stream of its own (the caller's may be the null stream)
    if (b->streams.empty()) {
        hipStream_t st;
        GT_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        b->streams.push_back(st);
    }
    hipStream_t cs = b->streams[0];
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
    GT_HIP(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
    fused_solve(b, cs);
    GT_HIP(hipStreamEndCapture(cs, &b->graph));
    GT_HIP(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
}

// Autotuner of the fused batch: NC2C table entries per work item, applied to
// every plan of the batch at once.  The per-plan rule (0) sizes items so that
// one problem alone gives two items per CU; in a fused launch all problems fill
// the chip together, and longer items mean fewer split tiles to reduce.  Each
// candidate's graph is replayed once to warm up and twice timed; the fastest
// is kept.  The choice is cached per problem set (matrix and grid sizes,
// samples, frames, coils, iterations) for the process and, where the
// environment variable GT_NUFFT_TUNE_CACHE names a file, across processes
// (one "signature length" line per set).
constexpr int kTuneLens[] = {0, 256, 512, 1024, 2048};

uint64_t fused_signature(const gt_cg_batch* b) {
    uint64_t h = 1469598103934665603ull;  // FNV-1a
    auto mix = [&](uint64_t v) {
        for (int k = 0; k < 8; k++) {
            h ^= (v >> (8 * k)) & 0xffu;
            h *= 1099511628211ull;
        }
    };
    mix(uint64_t(b->desc.iterations));
    for (const gt_sense_op* op : b->ops) {
        size_t Nos[4] = {0, 0, 0, 0};
        op->check(gt_nufft_get_matrix_size_os(op->plan, Nos));
        mix(op->N[0]);
        mix(op->N[1]);
        mix(Nos[0]);
        mix(Nos[1]);
        mix(op->M);
        mix(op->F);
        mix(op->C);
    }
    return h;
}

std::mutex tune_mu;
std::map<uint64_t, int>& tune_cache() {
    static std::map<uint64_t, int> cache;
    static bool loaded = false;
    if (!loaded) {
        loaded = true;
        if (const char* path = std::getenv("GT_NUFFT_TUNE_CACHE")) {
            std::ifstream in(path);
            std::string sig;
            int len = 0;
            while (in >> sig >> len) {
                char* end = nullptr;
                const uint64_t v = std::strtoull(sig.c_str(), &end, 16);
                if (end && *end == '\0') cache[v] = len;
            }
        }
    }
    return cache;
}

double time_graph(gt_cg_batch* b, hipStream_t s) {
    GT_HIP(hipGraphLaunch(b->exec, s));
    GT_HIP(hipStreamSynchronize(s));
    const auto t0 = std::chrono::steady_clock::now();
    GT_HIP(hipGraphLaunch(b->exec, s));
    GT_HIP(hipGraphLaunch(b->exec, s));
    GT_HIP(hipStreamSynchronize(s));
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

void fused_tune(gt_cg_batch* b, hipStream_t s) {
    const uint64_t sig = fused_signature(b);
    int chosen = -1;
    {
        std::lock_guard<std::mutex> lock(tune_mu);
        auto& c = tune_cache();
        auto it = c.find(sig);
        if (it != c.end()) chosen = it->second;
    }
    auto apply = [&](int len) {
        for (gt_sense_op* op : b->ops) plan_retile_nc2c(op->plan, len, s);
        fused_setup(b, s);
        b->fused_item_len = len;
    };
    if (chosen >= 0) {
        apply(chosen);
        return;
    }
    double best = 1e300;
    int best_len = 0;
    for (int len : kTuneLens) {
        apply(len);
        const double t = time_graph(b, s);
        if (t < best) {
            best = t;
            best_len = len;
        }
    }
    if (b->fused_item_len != best_len) apply(best_len);
    std::lock_guard<std::mutex> lock(tune_mu);
    tune_cache()[sig] = best_len;
    if (const char* path = std::getenv("GT_NUFFT_TUNE_CACHE")) {
        std::ofstream out(path, std::ios::app);
        out << std::hex << sig << std::dec << " " << best_len << "\n";
    }
}

}  // namespace

extern "C" {

gt_status gt_sense_create(gt_sense_op** out, gt_nufft_plan* plan, const void* csm, size_t ncoils, const float* dcw,
                          size_t dcw_len) {
    return guarded([&] {
        GT_REQUIRE(out && plan && csm, GT_ERR_INVALID_ARGUMENT, "null operator, plan or coil maps");
        *out = nullptr;
        GT_REQUIRE(ncoils > 0, GT_ERR_INVALID_ARGUMENT, "no coils");
        GT_REQUIRE(!dcw || dcw_len > 0, GT_ERR_INVALID_ARGUMENT, "dcw given with zero length");
        GT_REQUIRE(!plan_is_double(plan), GT_ERR_UNSUPPORTED, "the SENSE operator takes float (GT_FLOAT) plans");
        auto op = std::make_unique<gt_sense_op>();
        op->track();
        op->plan = plan;
        size_t N[4] = {0, 0, 0, 0};
        auto chk = [](gt_status s) {
            if (s != GT_OK) throw gt_error(s, last_error());
        };
        chk(gt_nufft_get_matrix_size(plan, N));
        chk(gt_nufft_get_num_samples_frames(plan, &op->M, &op->F));
        GT_REQUIRE(op->M > 0 && op->F > 0, GT_ERR_NOT_PREPROCESSED, "plan has not been preprocessed");
        int D = 0;
        while (D < 4 && N[D] > 0) D++;
        GT_REQUIRE(D >= 2 && D <= 3, GT_ERR_UNSUPPORTED, "the SENSE operator needs a 2D or 3D plan");
        op->D = D;
        op->nimg = 1;
        for (int d = 0; d < D; d++) {
            op->N[d] = N[d];
            op->nimg *= N[d];
        }
        op->C = ncoils;
        op->csm = static_cast<const float2*>(csm);
        op->dcw = dcw;
        op->dcw_len = dcw_len;
        GT_REQUIRE(op->nimg * op->F * op->C < (size_t(1) << 40), GT_ERR_UNSUPPORTED, "problem too large");
        GT_HIP(hipGetDevice(&op->device));
        op->coil_img.reserve(op->nimg * op->F * op->C * sizeof(float2));
        op->coil_smp.reserve(op->M * op->F * op->C * sizeof(float2));
        *out = op.release();
    });
}

gt_status gt_sense_destroy(gt_sense_op* op) {
    return guarded([&] { delete op; });
}

gt_status gt_sense_mult_M(gt_sense_op* op, const void* in, void* out, int accumulate, gt_stream_t stream) {
    return guarded([&] {
        GT_REQUIRE(op && in && out, GT_ERR_INVALID_ARGUMENT, "null argument");
        op->mult_M(in, out, accumulate != 0, S(stream));
    });
}

gt_status gt_sense_mult_MH(gt_sense_op* op, const void* in, void* out, int accumulate, gt_stream_t stream) {
    return guarded([&] {
        GT_REQUIRE(op && in && out, GT_ERR_INVALID_ARGUMENT, "null argument");
        op->mult_MH(in, out, accumulate != 0, S(stream));
    });
}

gt_status gt_sense_mult_MH_M(gt_sense_op* op, const void* in, void* out, int accumulate, gt_stream_t stream) {
    return guarded([&] {
        GT_REQUIRE(op && in && out, GT_ERR_INVALID_ARGUMENT, "null argument");
        op->mult_MH_M(in, out, accumulate != 0, S(stream));
    });
}

gt_status gt_cgsense_solve(gt_sense_op* op, const void* data, void* x, const gt_cg_desc* desc, double* rq_history,
                           gt_stream_t stream) {
    return guarded([&] {
        GT_REQUIRE(op && data && x, GT_ERR_INVALID_ARGUMENT, "null argument");
        validate_desc(desc);
        hipStream_t s = S(stream);
        op->reserve_cg(desc->iterations, desc->n_regs > 0);
        cg_solve(op, data, x, desc, false, s);
        if (rq_history) {
            GT_HIP(hipMemcpyAsync(rq_history, op->hist.ptr, size_t(desc->iterations + 1) * sizeof(double),
                                  hipMemcpyDeviceToHost, s));
            GT_HIP(hipStreamSynchronize(s));
        }
    });
}

gt_status gt_cgsense_solve_image(gt_sense_op* op, const void* rhs, void* x, const gt_cg_desc* desc,
                                 double* rq_history, gt_stream_t stream) {
    return guarded([&] {
        GT_REQUIRE(op && rhs && x, GT_ERR_INVALID_ARGUMENT, "null argument");
        validate_desc(desc);
        hipStream_t s = S(stream);
        op->reserve_cg(desc->iterations, desc->n_regs > 0);
        cg_solve(op, rhs, x, desc, false, s, true);
        if (rq_history) {
            GT_HIP(hipMemcpyAsync(rq_history, op->hist.ptr, size_t(desc->iterations + 1) * sizeof(double),
                                  hipMemcpyDeviceToHost, s));
            GT_HIP(hipStreamSynchronize(s));
        }
    });
}

gt_status gt_cgsense_batch_create(gt_cg_batch** out, gt_sense_op* const* ops, const void* const* data,
                                  void* const* x, int nprob, const gt_cg_desc* desc, int mode, gt_stream_t stream) {
    return guarded([&] {
        GT_REQUIRE(out && ops && data && x && nprob > 0, GT_ERR_INVALID_ARGUMENT, "null or empty problem set");
        *out = nullptr;
        validate_desc(desc);
        GT_REQUIRE(mode == GT_CG_PLAIN || mode == GT_CG_BATCHED || mode == GT_CG_FUSED, GT_ERR_INVALID_ARGUMENT,
                   "unknown batch mode");
        GT_REQUIRE(mode != GT_CG_FUSED || (!desc->precond && desc->n_regs == 0), GT_ERR_UNSUPPORTED,
                   "the fused batch takes an identity regulariser (kappa) only: no preconditioner or image "
                   "regularisers");
        // one descriptor serves problems of different sizes: no initial guess
        // or regulariser images (a preconditioner must divide every image)
        GT_REQUIRE(!desc->x0, GT_ERR_UNSUPPORTED, "a batch takes no initial guess");
        for (int k = 0; k < desc->n_regs; k++)
            GT_REQUIRE(!desc->regs[k].image, GT_ERR_UNSUPPORTED,
                       "a batch takes no regulariser images (one descriptor serves problems of different sizes)");
        auto b = std::make_unique<gt_cg_batch>();
        b->desc = *desc;
        b->regs.assign(desc->regs, desc->regs + desc->n_regs);
        b->desc.regs = b->regs.empty() ? nullptr : b->regs.data();
        b->mode = mode;
        GT_HIP(hipGetDevice(&b->device));
        hipStream_t s = S(stream);
        for (int i = 0; i < nprob; i++) {
            GT_REQUIRE(ops[i] && data[i] && x[i], GT_ERR_INVALID_ARGUMENT, "null problem entry");
            for (int j = 0; j < i; j++)
                GT_REQUIRE(ops[j] != ops[i] && ops[j]->plan != ops[i]->plan, GT_ERR_INVALID_ARGUMENT,
                           "problems of one batch need their own operator and plan (the branches run concurrently)");
            validate_sizes(ops[i], desc);
            ops[i]->reserve_cg(desc->iterations, desc->n_regs > 0);
            ops[i]->warm(s);
            b->ops.push_back(ops[i]);
            b->data.push_back(data[i]);
            b->x.push_back(x[i]);
        }
        if (mode == GT_CG_BATCHED) tune_graph(b.get(), s);
        if (mode == GT_CG_FUSED) fused_tune(b.get(), s);
        if (mode != GT_CG_PLAIN)
            for (auto* o : b->ops) b->gens.emplace_back(plan_generation(o->plan), o->generation);
        *out = b.release();
    });
}

gt_status gt_cgsense_batch_run(gt_cg_batch* b, gt_stream_t stream) {
    return guarded([&] {
        GT_REQUIRE(b, GT_ERR_INVALID_ARGUMENT, "null batch");
        hipStream_t s = S(stream);
        if (b->mode == GT_CG_BATCHED || b->mode == GT_CG_FUSED) {
            for (size_t i = 0; i < b->ops.size(); i++)
                GT_REQUIRE(b->gens[i].first == plan_generation(b->ops[i]->plan) &&
                               b->gens[i].second == b->ops[i]->generation,
                           GT_ERR_INVALID_ARGUMENT,
                           "a plan or operator bound to this batch changed after gt_cgsense_batch_create "
                           "(preprocess, set_tiling or a workspace reallocation): create the batch again");
            GT_HIP(hipGraphLaunch(b->exec, s));
        } else {
            for (size_t i = 0; i < b->ops.size(); i++) cg_solve(b->ops[i], b->data[i], b->x[i], &b->desc, true, s);
        }
    });
}

gt_status gt_image_operator_compute(const void* image, size_t n, int offset_estimation, float* offset, float* out,
                                    gt_stream_t stream) {
    return guarded([&] {
        GT_REQUIRE(image && offset && out, GT_ERR_INVALID_ARGUMENT, "null argument");
        GT_REQUIRE(n > 0 && n < (size_t(1) << 31), GT_ERR_INVALID_ARGUMENT, "image size must be in [1, 2^31)");
        hipStream_t s = S(stream);
        const long long nn = (long long)n;
        const unsigned nb = (unsigned)std::min<long long>(kRedBlocks, (nn + kRedThreads - 1) / kRedThreads);
        DevBuf part, scal, bmax, bins;
        part.reserve(nb * sizeof(double2));
        scal.reserve(4 * sizeof(float));
        bmax.reserve(nb * sizeof(float));
        const float2* x = static_cast<const float2*>(image);
        hipLaunchKernelGGL(asum_kernel, dim3(nb), dim3(kRedThreads), 0, s, x, nn, part.as<double2>());
        hipLaunchKernelGGL(asum_finish_kernel, dim3(1), dim3(kRedThreads), 0, s, part.as<double2>(), int(nb), nn,
                           scal.as<float>());
        hipLaunchKernelGGL(abs_scale_kernel, dim3(nb), dim3(kRedThreads), 0, s, x, nn, scal.as<float>(), out,
                           bmax.as<float>());
        hipLaunchKernelGGL(max_finish_kernel, dim3(1), dim3(kRedThreads), 0, s, bmax.as<float>(), int(nb),
                           scal.as<float>());
        GT_HIP(hipGetLastError());
        float h[4] = {0, 0, 0, 0};
        GT_HIP(hipMemcpyAsync(h, scal.ptr, 3 * sizeof(float), hipMemcpyDeviceToHost, s));
        GT_HIP(hipStreamSynchronize(s));
        GT_REQUIRE(h[2] > 0.0f && std::isfinite(h[0]), GT_ERR_INVALID_ARGUMENT,
                   "the regularisation image is zero (its normalisation divides by sum |x|)");
        if (offset_estimation) {
            // cuImageOperator::estimate_offset: histogram of image / max in
            // 50000 bins; the offset is (bins to reach 1 % of the elements + 1)
            // times max / 50000
            bins.reserve(kHistBins * sizeof(unsigned int));
            hipLaunchKernelGGL(zero_u32_kernel, dim3(blocks_for(kHistBins, 256)), dim3(256), 0, s,
                               bins.as<unsigned int>(), (long long)kHistBins);
            hipLaunchKernelGGL(hist_kernel, dim3(blocks_for(nn, 256)), dim3(256), 0, s, out, nn, scal.as<float>(),
                               bins.as<unsigned int>());
            GT_HIP(hipGetLastError());
            std::vector<unsigned int> hist(kHistBins);
            GT_HIP(hipMemcpyAsync(hist.data(), bins.ptr, kHistBins * sizeof(unsigned int), hipMemcpyDeviceToHost, s));
            GT_HIP(hipStreamSynchronize(s));
            const unsigned int target = (unsigned int)(0.01f * (float)n);
            unsigned int cumsum = 0, counter = 0;
            while (cumsum < target) cumsum += hist[counter++];
            *offset = float(counter + 1) * h[1] / float(kHistBins);
        }
        hipLaunchKernelGGL(offset_recip_kernel, dim3(blocks_for(nn, 256)), dim3(256), 0, s, out, nn, *offset);
        GT_HIP(hipGetLastError());
        GT_HIP(hipStreamSynchronize(s));  // the temporaries are freed on return
    });
}

gt_status gt_cgsense_precond_weights(const void* csm, size_t nimg, size_t ncoils, const float* reg, double kappa,
                                     void* weights, gt_stream_t stream) {
    return guarded([&] {
        GT_REQUIRE(csm && weights, GT_ERR_INVALID_ARGUMENT, "null argument");
        GT_REQUIRE(nimg > 0 && ncoils > 0, GT_ERR_INVALID_ARGUMENT, "empty image or no coils");
        hipLaunchKernelGGL(precond_weights_kernel, dim3(blocks_for((long long)nimg, 256)), dim3(256), 0, S(stream),
                           static_cast<const float2*>(csm), (long long)nimg, int(ncoils), reg, float(kappa),
                           static_cast<float2*>(weights));
        GT_HIP(hipGetLastError());
    });
}

gt_status gt_cgsense_batch_destroy(gt_cg_batch* b) {
    return guarded([&] { delete b; });
}

gt_status gt_cgsense_batch_branches(const gt_cg_batch* b, int* branches) {
    return guarded([&] {
        GT_REQUIRE(b && branches, GT_ERR_INVALID_ARGUMENT, "null argument");
        *branches = b->branches;
    });
}

gt_status gt_cgsense_batch_item_length(const gt_cg_batch* b, int* entries) {
    return guarded([&] {
        GT_REQUIRE(b && entries, GT_ERR_INVALID_ARGUMENT, "null argument");
        *entries = b->mode == GT_CG_FUSED ? b->fused_item_len : -1;
    });
}

}  // extern "C"
