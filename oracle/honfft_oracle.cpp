// honfft_oracle.cpp -- TEST INFRASTRUCTURE ONLY (the oracle).
//
// A CPU restatement of Gadgetron's hoNFFT path, written from the reference's
// algorithm.  It is the parity checker for the HIP library and the CPU
// baseline timed by bench.py.  Only tests/, __graft_entry__.smoke() and
// bench.py's cpu_baseline leg load it; the product path never links it.
//
// What it follows (reference file:line, relative to the gadgetron tree):
//   toolboxes/nfft/ConvolutionKernel.hpp:9-94     bessi0 (float and double forms), KaiserBessel
//   toolboxes/nfft/ConvolutionKernel.hpp:165-180  get(u): zero if any |u_d| > W/2; get(r, ax)
//   toolboxes/nfft/ConvolutionKernel.hpp:226-254  KaiserKernel ctors (os factor per dim)
//   toolboxes/nfft/ConvolutionKernel.hpp:287-303  compute_beta
//   toolboxes/nfft/cpu/ConvolutionMatrix.cpp:18-99   window ceil(p-r)..floor(p+r), wrap (i+N)%N,
//                                                    x fastest, per-sample row of (index, weight)
//   toolboxes/nfft/cpu/ConvolutionMatrix.cpp:102-137 transpose: per cell, samples ascending
//   toolboxes/nfft/cpu/hoGriddingConvolution.cpp:42-73   p = (k + 1/2) * N_os, one matrix per frame
//   toolboxes/nfft/cpu/hoGriddingConvolution.cpp:86-150  mvm; batch b uses matrix b % frames
//   toolboxes/nfft/cpu/hoNFFT.cpp:34-81     FFTD: centred FFT, do_scale=false multiplies by sqrt(N)
//   toolboxes/nfft/cpu/hoNFFT.cpp:84-197    deapodization filters (both ctor scalings)
//   toolboxes/nfft/NFFT.hpp:90-238          compute modes, pad/crop, dcw, op order, mult_MH_M
//   toolboxes/fft/cpu/hoNDFFT.cpp:130-180,576-586  ifftshift -> FFT -> fftshift, 1/sqrt(N)
//   toolboxes/core/cpu/hoNDArray_utils.h:579-650   pad offset N_os/2 - N/2
//
// Deliberate differences from the reference, each pinned by a test:
//   * FFTW is absent here.  The FFT below is a textbook mixed-radix Cooley-Tukey
//     (decimation in time, twiddles from double precision), checked against a
//     brute-force DFT, numpy.fft and Parseval.
//   * hoNFFT_plan::mult_MH_M reads sample/frame counts nothing assigns
//     (hoNFFT.cpp:243-244); this restatement uses NFFT.hpp:89-101 instead.
//   * Re-preprocessing clears the per-frame matrices (hoGriddingConvolution.cpp:58-72
//     appends without clearing).
//   * BACKWARDS_C2NC (NFFT.hpp:186-200) goes through compute_NFFTH_C2NC
//     (NFFT.hpp:232-238), which calls convolve(samples, image, NC2C); this
//     restatement computes the true adjoint of FORWARDS_NC2C instead.

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

thread_local std::string g_last_error;

// ---------------------------------------------------------------------------
// Kaiser-Bessel kernel (ConvolutionKernel.hpp:9-94).  The rational approximant
// of I0 is the reference's own; float and double forms are kept separate so
// the float oracle rounds like the float reference.
// ---------------------------------------------------------------------------
inline double bessi0(double x) {
    if (x == 0.0) return 1.0;
    const double z = x * x;
    const double num =
        (z * (z * (z * (z * (z * (z * (z * (z * (z * (z * (z * (z * (z *
            (z * 0.210580722890567e-22 + 0.380715242345326e-19) +
            0.479440257548300e-16) + 0.435125971262668e-13) +
            0.300931127112960e-10) + 0.160224679395361e-7) +
            0.654858370096785e-5) + 0.202591084143397e-2) +
            0.463076284721000e0) + 0.754337328948189e2) +
            0.830792541809429e4) + 0.571661130563785e6) +
            0.216415572361227e8) + 0.356644482244025e9) +
            0.144048298227235e10);
    const double den = (z * (z * (z - 0.307646912682801e4) +
                             0.347626332405882e7) - 0.144048298227235e10);
    return -num / den;
}

inline float bessi0(float x) {
    if (x == 0.0f) return 1.0f;
    const float z = x * x;
    const float num =
        (z * (z * (z * (z * (z * (z * (z * (z * (z * (z * (z * (z * (z *
            (z * 0.210580722890567e-22f + 0.380715242345326e-19f) +
            0.479440257548300e-16f) + 0.435125971262668e-13f) +
            0.300931127112960e-10f) + 0.160224679395361e-7f) +
            0.654858370096785e-5f) + 0.202591084143397e-2f) +
            0.463076284721000e0f) + 0.754337328948189e2f) +
            0.830792541809429e4f) + 0.571661130563785e6f) +
            0.216415572361227e8f) + 0.356644482244025e9f) +
            0.144048298227235e10f);
    const float den = (z * (z * (z - 0.307646912682801e4f) +
                            0.347626332405882e7f) - 0.144048298227235e10f);
    return -num / den;
}

template <class R>
inline R kaiser_bessel(R u, R nos, R one_over_W, R beta) {
    R t = R(2) * u * one_over_W;
    R tt = t * t;
    R arg = beta * std::sqrt(R(1) - tt);
    R b = bessi0(arg);
    return nos * b * one_over_W;
}

template <class R>
struct Kernel {
    int D = 2;
    R width = 0, radius = 0, width_inv = 0;
    R nos[4] = {0, 0, 0, 0};
    R os_factor[4] = {0, 0, 0, 0};
    R beta[4] = {0, 0, 0, 0};

    // ConvolutionKernel.hpp:287-303
    void compute_beta() {
        for (int d = 0; d < D; d++) {
            auto sqr = [](R x) { return x * x; };
            beta[d] = R(M_PI) * std::sqrt(sqr(width) / sqr(os_factor[d]) *
                                          sqr(os_factor[d] - R(0.5)) - R(0.8));
        }
    }
    // ConvolutionKernel.hpp:174-180 (1D form, used by the deapodization filter)
    R get1(R r, int ax) const {
        r = std::abs(r);
        if (r > radius) return R(0);
        return kaiser_bessel<R>(r, nos[ax], width_inv, beta[ax]);
    }
    // ConvolutionKernel.hpp:164-170 + :97-131 (separable product, x first)
    R get(const R* u) const {
        for (int d = 0; d < D; d++)
            if (u[d] > radius) return R(0);
        R v = kaiser_bessel<R>(u[0], nos[0], width_inv, beta[0]);
        for (int d = 1; d < D; d++) v = v * kaiser_bessel<R>(u[d], nos[d], width_inv, beta[d]);
        return v;
    }
};

// ---------------------------------------------------------------------------
// 1D complex FFT, mixed radix, recursive decimation in time.
// X[k] = sum_n x[n] exp(-+2 pi i k n / n_total); unnormalised.
// ---------------------------------------------------------------------------
template <class R>
class FFT1D {
public:
    using C = std::complex<R>;
    explicit FFT1D(size_t n) : n_(n) {
        size_t m = n;
        for (size_t p : {4u, 2u, 3u, 5u}) {
            while (m % p == 0) { factors_.push_back(p); m /= p; }
        }
        for (size_t p = 7; p * p <= m; p += 2) {
            while (m % p == 0) { factors_.push_back(p); m /= p; }
        }
        if (m > 1) factors_.push_back(m);
        tw_.resize(n);
        for (size_t k = 0; k < n; k++) {
            double a = -2.0 * M_PI * double(k) / double(n);
            tw_[k] = C(R(std::cos(a)), R(std::sin(a)));
        }
    }
    size_t size() const { return n_; }

    // in: strided input, out: contiguous output of length n.
    void exec(const C* in, size_t istride, C* out, bool forward) const {
        rec(in, istride, out, n_, 0, 1, forward);
    }

private:
    C w(size_t k, bool forward) const {
        const C& t = tw_[k % n_];
        return forward ? t : std::conj(t);
    }
    void rec(const C* in, size_t is, C* out, size_t n, size_t fi, size_t tws,
             bool fwd) const {
        if (n == 1) { out[0] = in[0]; return; }
        const size_t p = factors_[fi];
        const size_t m = n / p;
        for (size_t q = 0; q < p; q++)
            rec(in + q * is, is * p, out + q * m, m, fi + 1, tws * p, fwd);
        C z[64];
        std::vector<C> zbig;
        C* zp = z;
        if (p > 64) { zbig.resize(p); zp = zbig.data(); }
        const size_t np = n_ / p;  // twiddle stride of the p-point DFT
        for (size_t k = 0; k < m; k++) {
            zp[0] = out[k];
            for (size_t q = 1; q < p; q++) zp[q] = out[q * m + k] * w(q * k * tws, fwd);
            switch (p) {
            case 2: {
                out[k] = zp[0] + zp[1];
                out[k + m] = zp[0] - zp[1];
                break;
            }
            case 4: {
                C a0 = zp[0] + zp[2], a1 = zp[0] - zp[2];
                C b0 = zp[1] + zp[3], b1 = zp[1] - zp[3];
                // multiply b1 by -i (forward) or +i (backward)
                C jb1 = fwd ? C(b1.imag(), -b1.real()) : C(-b1.imag(), b1.real());
                out[k] = a0 + b0;
                out[k + m] = a1 + jb1;
                out[k + 2 * m] = a0 - b0;
                out[k + 3 * m] = a1 - jb1;
                break;
            }
            default: {
                C res[64];
                std::vector<C> rbig;
                C* rp = res;
                if (p > 64) { rbig.resize(p); rp = rbig.data(); }
                for (size_t t = 0; t < p; t++) {
                    C acc = zp[0];
                    for (size_t q = 1; q < p; q++) acc += zp[q] * w(((q * t) % p) * np, fwd);
                    rp[t] = acc;
                }
                for (size_t t = 0; t < p; t++) out[k + t * m] = rp[t];
            }
            }
        }
    }

    size_t n_;
    std::vector<size_t> factors_;
    std::vector<C> tw_;
};

// Unnormalised N-D FFT over the first `rank` dims of each batch block.
template <class R>
void fft_nd_unscaled(std::complex<R>* data, const std::vector<size_t>& dims, int rank,
                     bool forward) {
    using C = std::complex<R>;
    size_t block = 1;
    for (int d = 0; d < rank; d++) block *= dims[d];
    size_t total = 1;
    for (size_t v : dims) total *= v;
    const size_t nbatch = block ? total / block : 0;
    size_t stride = 1;
    for (int d = 0; d < rank; d++) {
        const size_t n = dims[d];
        FFT1D<R> plan(n);
        const size_t inner = stride;
        const size_t outer = block / (n * stride);
        const long long nlines = (long long)(nbatch * outer * inner);
#pragma omp parallel
        {
            std::vector<C> line(n), res(n);
#pragma omp for schedule(static)
            for (long long L = 0; L < nlines; L++) {
                const size_t b = size_t(L) / (outer * inner);
                const size_t rem = size_t(L) % (outer * inner);
                const size_t o = rem / inner, i = rem % inner;
                C* base = data + b * block + o * n * stride + i;
                for (size_t j = 0; j < n; j++) line[j] = base[j * stride];
                plan.exec(line.data(), 1, res.data(), forward);
                for (size_t j = 0; j < n; j++) base[j * stride] = res[j];
            }
        }
        stride *= n;
    }
}

// fftshift / ifftshift as rotations over the first `rank` dims
// (hoNDFFT.cpp:174-180: pivots (x+1)/2 and x-(x+1)/2; r[j] = a[(j+pivot) % x]).
template <class R>
void shift_nd(std::complex<R>* data, const std::vector<size_t>& dims, int rank, bool inverse) {
    size_t block = 1;
    for (int d = 0; d < rank; d++) block *= dims[d];
    size_t total = 1;
    for (size_t v : dims) total *= v;
    const size_t nbatch = block ? total / block : 0;
    std::vector<size_t> pivot(rank);
    for (int d = 0; d < rank; d++)
        pivot[d] = inverse ? dims[d] - (dims[d] + 1) / 2 : (dims[d] + 1) / 2;
#pragma omp parallel
    {
        std::vector<std::complex<R>> tmp(block);
        std::vector<size_t> co(rank);
#pragma omp for schedule(static)
        for (long long b = 0; b < (long long)nbatch; b++) {
            std::complex<R>* blk = data + size_t(b) * block;
            for (size_t idx = 0; idx < block; idx++) {
                size_t r = idx, src = 0, st = 1;
                for (int d = 0; d < rank; d++) {
                    size_t c = r % dims[d];
                    r /= dims[d];
                    src += ((c + pivot[d]) % dims[d]) * st;
                    st *= dims[d];
                }
                tmp[idx] = blk[src];
            }
            std::copy(tmp.begin(), tmp.end(), blk);
        }
    }
}

// hoNDFFT::fft{1,2,3}c / ifft{1,2,3}c: ifftshift -> FFT -> fftshift, times 1/sqrt(N).
template <class R>
void centred_fft(std::complex<R>* data, const std::vector<size_t>& dims, int rank, bool forward) {
    shift_nd<R>(data, dims, rank, true);
    fft_nd_unscaled<R>(data, dims, rank, forward);
    shift_nd<R>(data, dims, rank, false);
    size_t block = 1;
    for (int d = 0; d < rank; d++) block *= dims[d];
    size_t total = 1;
    for (size_t v : dims) total *= v;
    const R s = R(1) / std::sqrt(R(block));
#pragma omp parallel for schedule(static)
    for (long long i = 0; i < (long long)total; i++) data[i] *= s;
}

// hoNFFT.cpp:34-81 (FFTD): do_scale=false multiplies the orthonormal result by sqrt(N).
template <class R>
void fftd(std::complex<R>* data, const std::vector<size_t>& dims, int rank, bool forward,
          bool do_scale) {
    centred_fft<R>(data, dims, rank, forward);
    if (!do_scale) {
        size_t block = 1;
        for (int d = 0; d < rank; d++) block *= dims[d];
        size_t total = 1;
        for (size_t v : dims) total *= v;
        const R s = std::sqrt(R(block));
#pragma omp parallel for schedule(static)
        for (long long i = 0; i < (long long)total; i++) data[i] *= s;
    }
}

// ---------------------------------------------------------------------------
// Convolution matrices (ConvolutionMatrix.cpp:18-137).
// ---------------------------------------------------------------------------
template <class R>
struct ConvMatrix {
    size_t n_cols = 0;  // samples
    size_t n_rows = 0;  // grid cells
    std::vector<std::vector<uint32_t>> indices;
    std::vector<std::vector<R>> weights;
};

template <class R>
void window_rec(const Kernel<R>& k, const R* point, const size_t* nos, int D, int dim,
                R* image_point, size_t index, std::vector<uint32_t>& idx, std::vector<R>& wts) {
    if (dim < 0) {
        R delta[4];
        for (int d = 0; d < D; d++) delta[d] = std::abs(image_point[d] - point[d]);
        idx.push_back(uint32_t(index));
        wts.push_back(k.get(delta));
        return;
    }
    size_t frame_offset = 1;
    for (int d = 0; d < dim; d++) frame_offset *= nos[d];
    const int lo = int(std::ceil(point[dim] - k.radius));
    const int hi = int(std::floor(point[dim] + k.radius));
    for (int i = lo; i <= hi; i++) {
        const long long n = (long long)nos[dim];
        const size_t wrapped = size_t(((long long)i + n) % n);
        image_point[dim] = R(i);
        window_rec<R>(k, point, nos, D, dim - 1, image_point, index + frame_offset * wrapped, idx, wts);
    }
}

template <class R>
ConvMatrix<R> make_conv_matrix(const R* traj_scaled, size_t M, const size_t* nos, int D,
                               const Kernel<R>& k) {
    ConvMatrix<R> m;
    m.n_cols = M;
    m.n_rows = 1;
    for (int d = 0; d < D; d++) m.n_rows *= nos[d];
    m.indices.resize(M);
    m.weights.resize(M);
#pragma omp parallel for schedule(static)
    for (long long i = 0; i < (long long)M; i++) {
        R ip[4] = {0, 0, 0, 0};
        window_rec<R>(k, traj_scaled + size_t(i) * D, nos, D, D - 1, ip, 0, m.indices[i], m.weights[i]);
    }
    return m;
}

template <class R>
ConvMatrix<R> transpose(const ConvMatrix<R>& m) {
    ConvMatrix<R> t;
    t.n_cols = m.n_rows;
    t.n_rows = m.n_cols;
    t.indices.resize(m.n_rows);
    t.weights.resize(m.n_rows);
    std::vector<uint32_t> counts(m.n_rows, 0);
    for (size_t i = 0; i < m.n_cols; i++)
        for (auto r : m.indices[i]) counts[r]++;
    for (size_t r = 0; r < m.n_rows; r++) {
        t.indices[r].reserve(counts[r]);
        t.weights[r].reserve(counts[r]);
    }
    for (size_t i = 0; i < m.n_cols; i++) {
        for (size_t n = 0; n < m.indices[i].size(); n++) {
            t.indices[m.indices[i][n]].push_back(uint32_t(i));
            t.weights[m.indices[i][n]].push_back(m.weights[i][n]);
        }
    }
    return t;
}

// ---------------------------------------------------------------------------
// The plan (NFFT.hpp + hoNFFT.cpp + hoGriddingConvolution.cpp).
// ---------------------------------------------------------------------------
enum { FWD_C2NC = 0, FWD_NC2C = 1, BWD_C2NC = 2, BWD_NC2C = 3 };
enum { PREP_C2NC = 0, PREP_NC2C = 1, PREP_ALL = 2 };

struct PlanBase {
    virtual ~PlanBase() = default;
};

template <class R>
struct Plan : PlanBase {
    using C = std::complex<R>;
    int D = 2;
    size_t N[4] = {1, 1, 1, 1}, Nos[4] = {1, 1, 1, 1};
    R W = 0;
    Kernel<R> kernel;
    std::vector<C> deapod_ifft, deapod_fft;  // reciprocal filters, G elements
    size_t M = 0, F = 0;
    std::vector<ConvMatrix<R>> cm, cmT;

    size_t G() const { size_t g = 1; for (int d = 0; d < D; d++) g *= Nos[d]; return g; }
    size_t Nimg() const { size_t g = 1; for (int d = 0; d < D; d++) g *= N[d]; return g; }

    // hoNFFT.cpp:84-145 (separable grid of 1D get(x - N_os/2, ax))
    std::vector<C> filter_grid() const {
        std::vector<C> f(G());
        const size_t g = G();
        for (size_t idx = 0; idx < g; idx++) {
            size_t r = idx;
            R w = R(1);
            R ws[4] = {R(1), R(1), R(1), R(1)};
            for (int d = 0; d < D; d++) {
                size_t c = r % Nos[d];
                r /= Nos[d];
                R off = R(int(c)) - R(Nos[d]) / R(2);
                ws[d] = kernel.get1(off, d);
            }
            // hoNFFT multiplies weight_x*weight_y(*weight_z) in that order
            w = ws[0];
            for (int d = 1; d < D; d++) w = w * ws[d];
            f[idx] = C(w, R(0));
        }
        return f;
    }

    // hoNFFT.cpp:150-197; do_scale selects the ctor convention.
    void build_filters(bool do_scale) {
        std::vector<size_t> dims(Nos, Nos + D);
        deapod_ifft = filter_grid();
        deapod_fft = deapod_ifft;
        fftd<R>(deapod_ifft.data(), dims, D, false, do_scale);
        fftd<R>(deapod_fft.data(), dims, D, true, do_scale);
        for (auto& v : deapod_ifft) v = R(1) / v;
        for (auto& v : deapod_fft) v = R(1) / v;
    }

    // hoGriddingConvolution.cpp:42-73
    void preprocess(const R* traj, size_t M_, size_t F_, int mode) {
        M = M_;
        F = F_;
        cm.clear();
        cmT.clear();
        std::vector<R> scaled(M * F * D);
        for (size_t i = 0; i < M * F; i++)
            for (int d = 0; d < D; d++)
                scaled[i * D + d] = (traj[i * D + d] + R(0.5)) * R(Nos[d]);
        for (size_t f = 0; f < F; f++) {
            cm.push_back(make_conv_matrix<R>(scaled.data() + f * M * D, M, Nos, D, kernel));
            if (mode == PREP_NC2C || mode == PREP_ALL) cmT.push_back(transpose(cm.back()));
        }
    }

    // hoGriddingConvolution.cpp:86-150
    void conv_c2nc(const C* image, C* samples, size_t nbatches, bool accumulate) const {
        if (!accumulate) std::fill(samples, samples + nbatches * M, C(0));
#pragma omp parallel for schedule(dynamic)
        for (long long b = 0; b < (long long)nbatches; b++) {
            const ConvMatrix<R>& m = cm[size_t(b) % cm.size()];
            const C* img = image + size_t(b) * m.n_rows;
            C* s = samples + size_t(b) * m.n_cols;
            for (size_t i = 0; i < m.n_cols; i++) {
                const auto& ri = m.indices[i];
                const auto& wi = m.weights[i];
                for (size_t n = 0; n < ri.size(); n++) s[i] += img[ri[n]] * wi[n];
            }
        }
    }
    void conv_nc2c(const C* samples, C* image, size_t nbatches, bool accumulate) const {
        const size_t g = G();
        if (!accumulate) std::fill(image, image + nbatches * g, C(0));
#pragma omp parallel for schedule(dynamic)
        for (long long b = 0; b < (long long)nbatches; b++) {
            const ConvMatrix<R>& m = cmT[size_t(b) % cmT.size()];
            C* img = image + size_t(b) * m.n_cols;
            const C* s = samples + size_t(b) * m.n_rows;
            for (size_t r = 0; r < m.n_cols; r++) {
                const auto& ri = m.indices[r];
                const auto& wi = m.weights[r];
                for (size_t n = 0; n < ri.size(); n++) img[r] += s[ri[n]] * wi[n];
            }
        }
    }

    void fft(C* data, size_t nbatches, bool forward, bool do_scale) const {
        std::vector<size_t> dims(Nos, Nos + D);
        dims.push_back(nbatches);
        fftd<R>(data, dims, D, forward, do_scale);
    }
    void deapodize(C* data, size_t nbatches, bool fourier_domain) const {
        const std::vector<C>& f = fourier_domain ? deapod_fft : deapod_ifft;
        const size_t g = G();
#pragma omp parallel for schedule(static)
        for (long long i = 0; i < (long long)(g * nbatches); i++) data[i] *= f[size_t(i) % g];
    }

    // hoNDArray_utils.h:579-650 / cuNDArray_utils.cu:283-286: offset N_os/2 - N/2
    void pad(const C* in, C* out, size_t nbatches) const {
        const size_t g = G(), n = Nimg();
        std::fill(out, out + g * nbatches, C(0));
        size_t off[4];
        for (int d = 0; d < D; d++) off[d] = Nos[d] / 2 - N[d] / 2;
        for (size_t b = 0; b < nbatches; b++)
            for (size_t idx = 0; idx < n; idx++) {
                size_t r = idx, dst = 0, st = 1;
                for (int d = 0; d < D; d++) {
                    size_t c = r % N[d];
                    r /= N[d];
                    dst += (c + off[d]) * st;
                    st *= Nos[d];
                }
                out[b * g + dst] = in[b * n + idx];
            }
    }
    // NFFT.hpp:155,178: crop offset (N_os - N) >> 1
    void crop(const C* in, C* out, size_t nbatches) const {
        const size_t g = G(), n = Nimg();
        size_t off[4];
        for (int d = 0; d < D; d++) off[d] = (Nos[d] - N[d]) >> 1;
        for (size_t b = 0; b < nbatches; b++)
            for (size_t idx = 0; idx < n; idx++) {
                size_t r = idx, src = 0, st = 1;
                for (int d = 0; d < D; d++) {
                    size_t c = r % N[d];
                    r /= N[d];
                    src += (c + off[d]) * st;
                    st *= Nos[d];
                }
                out[b * n + idx] = in[b * g + src];
            }
    }

    static void mul_dcw(C* s, size_t n, const R* dcw, size_t dcw_len) {
        // cuNDArray_operators.cu:129-145 broadcast by modulus
        for (size_t i = 0; i < n; i++) s[i] *= dcw[i % dcw_len];
    }

    // NFFT.hpp:104-202
    void compute(const C* in, const std::vector<size_t>& in_dims, C* out,
                 const std::vector<size_t>& out_dims, const R* dcw, size_t dcw_len, int mode) {
        if (cm.empty()) throw std::runtime_error("compute: plan not preprocessed");
        const bool image_is_in = (mode == FWD_C2NC || mode == BWD_C2NC);
        const std::vector<size_t>& idims = image_is_in ? in_dims : out_dims;
        if (idims.size() < size_t(D)) throw std::runtime_error("compute: image rank < D");
        bool oversampled = true, plain = true;
        for (int d = 0; d < D; d++) {
            if (idims[d] != Nos[d]) oversampled = false;
            if (idims[d] != N[d]) plain = false;
        }
        if (!oversampled && !plain) throw std::runtime_error("compute: image dims match neither N nor N_os");
        size_t img_elems = 1;
        for (size_t v : idims) img_elems *= v;
        const size_t nb = img_elems / (oversampled ? G() : Nimg());  // batches incl. frames
        const size_t g = G();
        const size_t ns = nb * M;  // sample elements (nb counts frames too)
        std::vector<C> work(g * nb);
        switch (mode) {
        case FWD_C2NC: {
            if (!oversampled) pad(in, work.data(), nb);
            else std::copy(in, in + g * nb, work.begin());
            deapodize(work.data(), nb, false);
            fft(work.data(), nb, true, true);
            conv_c2nc(work.data(), out, nb, false);
            if (dcw) mul_dcw(out, ns, dcw, dcw_len);
            break;
        }
        case FWD_NC2C: {
            std::vector<C> s(in, in + ns);
            if (dcw) mul_dcw(s.data(), ns, dcw, dcw_len);
            C* dst = oversampled ? out : work.data();
            conv_nc2c(s.data(), dst, nb, false);
            fft(dst, nb, true, true);
            deapodize(dst, nb, true);
            if (!oversampled) crop(work.data(), out, nb);
            break;
        }
        case BWD_NC2C: {
            std::vector<C> s(in, in + ns);
            if (dcw) mul_dcw(s.data(), ns, dcw, dcw_len);
            C* dst = oversampled ? out : work.data();
            conv_nc2c(s.data(), dst, nb, false);
            fft(dst, nb, false, true);
            deapodize(dst, nb, false);
            if (!oversampled) crop(work.data(), out, nb);
            break;
        }
        case BWD_C2NC: {
            // adjoint of FORWARDS_NC2C: pad -> deapodize(fourier) -> FFT backwards -> C2NC -> dcw
            if (!oversampled) pad(in, work.data(), nb);
            else std::copy(in, in + g * nb, work.begin());
            deapodize(work.data(), nb, true);
            fft(work.data(), nb, false, true);
            conv_c2nc(work.data(), out, nb, false);
            if (dcw) mul_dcw(out, ns, dcw, dcw_len);
            break;
        }
        default: throw std::runtime_error("compute: unknown mode");
        }
    }
};

std::unique_ptr<PlanBase> make_plan(int D, int precision, const size_t* N, const size_t* Nos,
                                    double W, int ctor, double alpha) {
    if (D < 1 || D > 4) throw std::runtime_error("D must be 1 to 4");
    auto fill = [&](auto* p) {
        using R = std::remove_reference_t<decltype(p->W)>;
        p->D = D;
        p->W = R(W);
        p->kernel.D = D;
        p->kernel.width = R(W);
        p->kernel.radius = R(W) / R(2);
        p->kernel.width_inv = R(1) / R(W);
        for (int d = 0; d < D; d++) {
            p->N[d] = N[d];
            if (ctor == 1) {
                // NFFT.hpp:23-37 / ConvolutionKernel.hpp:241-254
                R a = R(alpha);
                p->Nos[d] = size_t(a * R(N[d]));
                p->kernel.os_factor[d] = a;
            } else {
                // ConvolutionKernel.hpp:225-238
                p->Nos[d] = Nos[d];
                p->kernel.os_factor[d] = R(double(unsigned(Nos[d]))) / R(double(unsigned(N[d])));
            }
            p->kernel.nos[d] = R(double(unsigned(p->Nos[d])));
            if (p->Nos[d] < p->N[d]) throw std::runtime_error("N_os must be >= N");
        }
        p->kernel.compute_beta();
        // ctor 0 (hoNFFT(N, N_os, W)) builds filters with do_scale=true,
        // ctor 1 (hoNFFT(N, alpha, W), the cuNFFT convention) with do_scale=false,
        // ctor 2 (cuNFFT_plan(N, N_os, W): N_os given, the unscaled filter FFT of
        // cuNFFT.cu:170-202) with do_scale=false as well.
        p->build_filters(ctor == 0);
    };
    if (precision == 0) {
        auto p = std::make_unique<Plan<float>>();
        fill(p.get());
        return p;
    }
    auto p = std::make_unique<Plan<double>>();
    fill(p.get());
    return p;
}

template <class F>
int guarded(F&& f) {
    try {
        f();
        return 0;
    } catch (const std::exception& e) {
        g_last_error = e.what();
        return 1;
    }
}

std::vector<size_t> to_dims(const size_t* d, int n) { return std::vector<size_t>(d, d + n); }

}  // namespace

// ---------------------------------------------------------------------------
// C API used by tests/ and bench.py through ctypes.
// ---------------------------------------------------------------------------
extern "C" {

struct gto_plan {
    std::unique_ptr<PlanBase> p;
    int precision;
};

const char* gto_last_error() { return g_last_error.c_str(); }

int gto_num_threads() {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void gto_set_num_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

// ctor: 0 = hoNFFT_plan(N, N_os, W)   (do_scale=true filters)
//       1 = hoNFFT_plan(N, alpha, W)  (do_scale=false filters; cuNFFT convention)
gto_plan* gto_plan_create(int D, int precision, const size_t* N, const size_t* Nos, double W,
                          int ctor, double alpha) {
    gto_plan* h = nullptr;
    int rc = guarded([&] {
        auto p = make_plan(D, precision, N, Nos, W, ctor, alpha);
        h = new gto_plan{std::move(p), precision};
    });
    return rc ? nullptr : h;
}

void gto_plan_destroy(gto_plan* h) { delete h; }

int gto_plan_info(gto_plan* h, size_t* Nos_out, double* beta_out) {
    return guarded([&] {
        if (h->precision == 0) {
            auto* p = static_cast<Plan<float>*>(h->p.get());
            for (int d = 0; d < p->D; d++) { Nos_out[d] = p->Nos[d]; beta_out[d] = p->kernel.beta[d]; }
        } else {
            auto* p = static_cast<Plan<double>*>(h->p.get());
            for (int d = 0; d < p->D; d++) { Nos_out[d] = p->Nos[d]; beta_out[d] = p->kernel.beta[d]; }
        }
    });
}

int gto_preprocess(gto_plan* h, const void* traj, size_t M, size_t F, int mode) {
    return guarded([&] {
        if (h->precision == 0) static_cast<Plan<float>*>(h->p.get())->preprocess((const float*)traj, M, F, mode);
        else static_cast<Plan<double>*>(h->p.get())->preprocess((const double*)traj, M, F, mode);
    });
}

int gto_compute(gto_plan* h, const void* in, const size_t* in_dims, int in_ndim, void* out,
                const size_t* out_dims, int out_ndim, const void* dcw, size_t dcw_len, int mode) {
    return guarded([&] {
        auto id = to_dims(in_dims, in_ndim), od = to_dims(out_dims, out_ndim);
        if (h->precision == 0) {
            auto* p = static_cast<Plan<float>*>(h->p.get());
            p->compute((const std::complex<float>*)in, id, (std::complex<float>*)out, od,
                       (const float*)dcw, dcw_len, mode);
        } else {
            auto* p = static_cast<Plan<double>*>(h->p.get());
            p->compute((const std::complex<double>*)in, id, (std::complex<double>*)out, od,
                       (const double*)dcw, dcw_len, mode);
        }
    });
}

// NFFT.hpp:89-101: FORWARDS_C2NC into [M, F, batches], then BACKWARDS_NC2C.
int gto_mult_MH_M(gto_plan* h, const void* in, const size_t* dims, int ndim, void* out,
                  const void* dcw, size_t dcw_len) {
    return guarded([&] {
        auto d = to_dims(dims, ndim);
        auto run = [&](auto* p, auto* inp, auto* outp, auto* dc) {
            using C = std::remove_const_t<std::remove_pointer_t<decltype(inp)>>;
            size_t img = 1;
            for (size_t v : d) img *= v;
            const size_t nb = img / p->Nimg();
            std::vector<C> tmp(p->M * nb);
            std::vector<size_t> sd = {p->M * nb};
            p->compute(inp, d, tmp.data(), sd, dc, dcw_len, FWD_C2NC);
            p->compute(tmp.data(), sd, outp, d, dc, dcw_len, BWD_NC2C);
        };
        if (h->precision == 0)
            run(static_cast<Plan<float>*>(h->p.get()), (const std::complex<float>*)in,
                (std::complex<float>*)out, (const float*)dcw);
        else
            run(static_cast<Plan<double>*>(h->p.get()), (const std::complex<double>*)in,
                (std::complex<double>*)out, (const double*)dcw);
    });
}

// NFFT.hpp:65-86 (standalone convolution on an oversampled image)
int gto_convolve(gto_plan* h, const void* in, void* out, size_t nbatches, int conv_mode,
                 int accumulate) {
    return guarded([&] {
        auto run = [&](auto* p, auto* inp, auto* outp) {
            if (conv_mode == 0) p->conv_c2nc(inp, outp, nbatches, accumulate != 0);
            else p->conv_nc2c(inp, outp, nbatches, accumulate != 0);
        };
        if (h->precision == 0)
            run(static_cast<Plan<float>*>(h->p.get()), (const std::complex<float>*)in,
                (std::complex<float>*)out);
        else
            run(static_cast<Plan<double>*>(h->p.get()), (const std::complex<double>*)in,
                (std::complex<double>*)out);
    });
}

int gto_fft(gto_plan* h, void* data, size_t nbatches, int forward, int do_scale) {
    return guarded([&] {
        if (h->precision == 0)
            static_cast<Plan<float>*>(h->p.get())->fft((std::complex<float>*)data, nbatches, forward != 0, do_scale != 0);
        else
            static_cast<Plan<double>*>(h->p.get())->fft((std::complex<double>*)data, nbatches, forward != 0, do_scale != 0);
    });
}

int gto_deapodize(gto_plan* h, void* data, size_t nbatches, int fourier_domain) {
    return guarded([&] {
        if (h->precision == 0)
            static_cast<Plan<float>*>(h->p.get())->deapodize((std::complex<float>*)data, nbatches, fourier_domain != 0);
        else
            static_cast<Plan<double>*>(h->p.get())->deapodize((std::complex<double>*)data, nbatches, fourier_domain != 0);
    });
}

// Reciprocal deapodization filter (G complex values) as the plan holds it.
int gto_get_deapod_filter(gto_plan* h, void* out, int fourier_domain) {
    return guarded([&] {
        if (h->precision == 0) {
            auto* p = static_cast<Plan<float>*>(h->p.get());
            const auto& f = fourier_domain ? p->deapod_fft : p->deapod_ifft;
            std::memcpy(out, f.data(), f.size() * sizeof(f[0]));
        } else {
            auto* p = static_cast<Plan<double>*>(h->p.get());
            const auto& f = fourier_domain ? p->deapod_fft : p->deapod_ifft;
            std::memcpy(out, f.data(), f.size() * sizeof(f[0]));
        }
    });
}

// Convolution matrix rows for frame 0 (for closed-form / invariant tests).
int gto_conv_row(gto_plan* h, size_t sample, uint32_t* idx_out, void* w_out, int cap, int* n_out) {
    return guarded([&] {
        auto get = [&](auto* p) {
            if (p->cm.empty()) throw std::runtime_error("not preprocessed");
            const auto& ri = p->cm[0].indices.at(sample);
            const auto& wi = p->cm[0].weights.at(sample);
            *n_out = int(ri.size());
            using R = std::remove_reference_t<decltype(wi[0])>;
            for (int n = 0; n < int(ri.size()) && n < cap; n++) {
                idx_out[n] = ri[n];
                ((std::remove_const_t<R>*)w_out)[n] = wi[n];
            }
        };
        if (h->precision == 0) get(static_cast<Plan<float>*>(h->p.get()));
        else get(static_cast<Plan<double>*>(h->p.get()));
    });
}

// hoNDFFT fft{1,2,3}c / ifft{1,2,3}c over the first `rank` dims (orthonormal).
int gto_centred_fft(int precision, void* data, const size_t* dims, int ndim, int rank, int forward) {
    return guarded([&] {
        auto d = to_dims(dims, ndim);
        if (precision == 0) centred_fft<float>((std::complex<float>*)data, d, rank, forward != 0);
        else centred_fft<double>((std::complex<double>*)data, d, rank, forward != 0);
    });
}

// Unnormalised, unshifted FFT (for brute-force DFT checks).
int gto_plain_fft(int precision, void* data, const size_t* dims, int ndim, int rank, int forward) {
    return guarded([&] {
        auto d = to_dims(dims, ndim);
        if (precision == 0) fft_nd_unscaled<float>((std::complex<float>*)data, d, rank, forward != 0);
        else fft_nd_unscaled<double>((std::complex<double>*)data, d, rank, forward != 0);
    });
}

int gto_fftshift(int precision, void* data, const size_t* dims, int ndim, int rank, int inverse) {
    return guarded([&] {
        auto d = to_dims(dims, ndim);
        if (precision == 0) shift_nd<float>((std::complex<float>*)data, d, rank, inverse != 0);
        else shift_nd<double>((std::complex<double>*)data, d, rank, inverse != 0);
    });
}

double gto_bessi0(double x) { return bessi0(x); }
float gto_bessi0f(float x) { return bessi0(x); }
double gto_kaiser_bessel(double u, double nos, double one_over_W, double beta) {
    return kaiser_bessel<double>(u, nos, one_over_W, beta);
}
float gto_kaiser_besself(float u, float nos, float one_over_W, float beta) {
    return kaiser_bessel<float>(u, nos, one_over_W, beta);
}

// FP64 direct NDFT: forward y_j = sum_n x_n exp(-2 pi i k_j . (n - N/2)),
// adjoint x_n = sum_j y_j exp(+2 pi i k_j . (n - N/2)).  Batched over nb images.
int gto_ndft(int D, const size_t* N, const double* traj, size_t M, const void* in, void* out,
             size_t nb, int adjoint) {
    return guarded([&] {
        using C = std::complex<double>;
        size_t n = 1;
        for (int d = 0; d < D; d++) n *= N[d];
        const C* x = (const C*)in;
        C* y = (C*)out;
        if (!adjoint) {
#pragma omp parallel for schedule(static)
            for (long long j = 0; j < (long long)M; j++) {
                const double* k = traj + size_t(j) * D;
                // separable phase tables
                std::vector<std::vector<C>> ph(D);
                for (int d = 0; d < D; d++) {
                    ph[d].resize(N[d]);
                    for (size_t c = 0; c < N[d]; c++) {
                        double a = -2.0 * M_PI * k[d] * (double(c) - double(N[d] / 2));
                        ph[d][c] = C(std::cos(a), std::sin(a));
                    }
                }
                for (size_t b = 0; b < nb; b++) {
                    C acc(0);
                    for (size_t idx = 0; idx < n; idx++) {
                        size_t r = idx;
                        C p(1);
                        for (int d = 0; d < D; d++) { p *= ph[d][r % N[d]]; r /= N[d]; }
                        acc += x[b * n + idx] * p;
                    }
                    y[b * M + size_t(j)] = acc;
                }
            }
        } else {
            const C* yy = (const C*)in;
            C* xx = (C*)out;
#pragma omp parallel for schedule(static)
            for (long long idx = 0; idx < (long long)n; idx++) {
                size_t r = size_t(idx);
                size_t co[4] = {0, 0, 0, 0};
                for (int d = 0; d < D; d++) { co[d] = r % N[d]; r /= N[d]; }
                for (size_t b = 0; b < nb; b++) {
                    C acc(0);
                    for (size_t j = 0; j < M; j++) {
                        double a = 0;
                        for (int d = 0; d < D; d++)
                            a += traj[j * D + d] * (double(co[d]) - double(N[d] / 2));
                        a *= 2.0 * M_PI;
                        acc += yy[b * M + j] * C(std::cos(a), std::sin(a));
                    }
                    xx[b * n + size_t(idx)] = acc;
                }
            }
        }
    });
}

}  // extern "C"
