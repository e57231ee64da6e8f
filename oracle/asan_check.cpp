// TEST INFRASTRUCTURE ONLY: drives the oracle (honfft_oracle.cpp) through its
// C API under AddressSanitizer and UndefinedBehaviorSanitizer (built by
// `make -C oracle asan`, run by tests/test_oracle_asan.py).  Every plan kind
// (D = 1..4, float and double, both constructors) runs every entry point on
// small random inputs, edge sizes included; the sanitizers turn any
// out-of-bounds access, leak or undefined operation into a failing exit.
#include <complex>
#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <random>
#include <vector>

extern "C" {
struct gto_plan;
const char* gto_last_error();
void gto_set_num_threads(int n);
gto_plan* gto_plan_create(int D, int precision, const size_t* N, const size_t* Nos, double W, int ctor, double alpha);
void gto_plan_destroy(gto_plan* h);
int gto_preprocess(gto_plan* h, const void* traj, size_t M, size_t F, int mode);
int gto_compute(gto_plan* h, const void* in, const size_t* in_dims, int in_ndim, void* out, const size_t* out_dims,
                int out_ndim, const void* dcw, size_t dcw_len, int mode);
int gto_mult_MH_M(gto_plan* h, const void* in, const size_t* dims, int ndim, void* out, const void* dcw,
                  size_t dcw_len);
int gto_convolve(gto_plan* h, const void* in, void* out, size_t nbatches, int conv_mode, int accumulate);
int gto_fft(gto_plan* h, void* data, size_t nbatches, int forward, int do_scale);
int gto_deapodize(gto_plan* h, void* data, size_t nbatches, int fourier_domain);
int gto_get_deapod_filter(gto_plan* h, void* out, int fourier_domain);
int gto_ndft(int D, const size_t* N, const double* traj, size_t M, const void* in, void* out, size_t nb, int adjoint);
}

namespace {

int failures = 0;

void expect(int rc, const char* what) {
    if (rc != 0) {
        std::fprintf(stderr, "asan_check: %s failed: %s\n", what, gto_last_error());
        failures++;
    }
}

template <class R>
void run_case(int D, const size_t* N, int ctor, size_t M, size_t F, size_t C, std::mt19937& rng) {
    using Cx = std::complex<R>;
    const int prec = sizeof(R) == 4 ? 0 : 1;
    size_t Nos[4] = {0, 0, 0, 0};
    for (int d = 0; d < D; d++) Nos[d] = (N[d] * 3 + 1) / 2 + ((N[d] * 3 + 1) / 2) % 2;
    gto_plan* p = gto_plan_create(D, prec, N, Nos, 5.5, ctor, 1.5);
    if (!p) {
        std::fprintf(stderr, "asan_check: plan create failed: %s\n", gto_last_error());
        failures++;
        return;
    }
    std::uniform_real_distribution<double> U(-0.5, 0.5);
    std::normal_distribution<double> G(0.0, 1.0);
    std::vector<R> traj(M * F * D);
    for (auto& v : traj) v = R(U(rng));
    if (M) traj[0] = R(-0.5);  // the edges of the range
    if (M > 1) traj[D] = R(0.5 - 1e-7);
    expect(gto_preprocess(p, traj.data(), M, F, 2), "preprocess");
    size_t nimg = 1, nos = 1;
    for (int d = 0; d < D; d++) {
        nimg *= N[d];
        nos *= Nos[d];
    }
    (void)nos;
    std::vector<Cx> img(nimg * F * C), smp(M * F * C), out_s(M * F * C), out_i(nimg * F * C);
    for (auto& v : img) v = Cx(R(G(rng)), R(G(rng)));
    for (auto& v : smp) v = Cx(R(G(rng)), R(G(rng)));
    std::vector<R> dcw(M * F);
    for (auto& v : dcw) v = R(U(rng) + 0.5);
    std::vector<size_t> idims(N, N + D), sdims = {M, F, C};
    idims.push_back(F);
    idims.push_back(C);
    for (int mode = 0; mode < 4; mode++) {
        const bool image_in = mode == 0 || mode == 2;
        for (int with_dcw = 0; with_dcw < 2; with_dcw++) {
            const void* w = with_dcw ? dcw.data() : nullptr;
            if (image_in)
                expect(gto_compute(p, img.data(), idims.data(), int(idims.size()), out_s.data(), sdims.data(), 3, w,
                                   with_dcw ? dcw.size() : 0, mode),
                       "compute (image in)");
            else
                expect(gto_compute(p, smp.data(), sdims.data(), 3, out_i.data(), idims.data(), int(idims.size()), w,
                                   with_dcw ? dcw.size() : 0, mode),
                       "compute (samples in)");
        }
    }
    expect(gto_mult_MH_M(p, img.data(), idims.data(), int(idims.size()), out_i.data(), dcw.data(), dcw.size()),
           "mult_MH_M");
    size_t G_os = 1;
    for (int d = 0; d < D; d++) G_os *= Nos[d];
    std::vector<Cx> grid(G_os * F * C);
    for (auto& v : grid) v = Cx(R(G(rng)), R(G(rng)));
    expect(gto_convolve(p, grid.data(), out_s.data(), F * C, 0, 0), "convolve C2NC");
    expect(gto_convolve(p, smp.data(), grid.data(), F * C, 1, 1), "convolve NC2C accumulate");
    expect(gto_fft(p, grid.data(), F * C, 1, 1), "fft forwards");
    expect(gto_fft(p, grid.data(), F * C, 0, 0), "fft backwards");
    expect(gto_deapodize(p, grid.data(), F * C, 0), "deapodize");
    expect(gto_deapodize(p, grid.data(), F * C, 1), "deapodize fourier");
    std::vector<Cx> filt(G_os);
    expect(gto_get_deapod_filter(p, filt.data(), 0), "deapod filter");
    gto_plan_destroy(p);
}

}  // namespace

int main() {
    gto_set_num_threads(2);
    std::mt19937 rng(7);
    const size_t n1[1] = {16}, n2[2] = {12, 10}, n3[3] = {6, 8, 6}, n4[4] = {4, 4, 4, 4};
    int cases = 0;
    for (int ctor = 0; ctor < 3; ctor++) {
        run_case<float>(1, n1, ctor, 40, 2, 3, rng);
        run_case<double>(1, n1, ctor, 40, 1, 2, rng);
        run_case<float>(2, n2, ctor, 60, 2, 2, rng);
        run_case<double>(2, n2, ctor, 60, 1, 1, rng);
        run_case<float>(3, n3, ctor, 30, 1, 2, rng);
        run_case<double>(3, n3, ctor, 30, 1, 1, rng);
        run_case<float>(4, n4, ctor, 12, 1, 1, rng);
        cases += 7;
    }
    run_case<float>(2, n2, 1, 1, 1, 1, rng);  // a single sample
    cases++;
    // the FP64 NDFT in both directions
    std::vector<double> k(2 * 20);
    std::uniform_real_distribution<double> U(-0.5, 0.5);
    for (auto& v : k) v = U(rng);
    std::vector<std::complex<double>> x(12 * 10 * 2, {1.0, -0.5}), y(20 * 2);
    expect(gto_ndft(2, n2, k.data(), 20, x.data(), y.data(), 2, 0), "ndft");
    expect(gto_ndft(2, n2, k.data(), 20, y.data(), x.data(), 2, 1), "ndft adjoint");
    cases++;
    if (failures) {
        std::fprintf(stderr, "asan_check: %d failures\n", failures);
        return 1;
    }
    std::printf("asan_check: %d plan cases ok\n", cases);
    return 0;
}
