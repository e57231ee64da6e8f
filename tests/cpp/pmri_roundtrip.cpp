// pmri_roundtrip.cpp -- drives libgtnufft.so through the C++ front end of the
// parallel-imaging calls (include/gt_pmri.hpp) in the order of the reference's
// golden-angle CG-SENSE application (apps/standalone/gpu/mri/sense/
// noncartesian/radial/2d_golden_ratio/main_cg.cpp:100-200): trajectory and
// DCW, frames into a cuSenseBuffer, coil maps from its images, then the SENSE
// operator and CG on one frame.
//
// usage: pmri_roundtrip <dir>
//   reads  <dir>/meta.txt      "N C S P F"  (N x N image, C coils, S samples per
//                              profile, P profiles per frame, F frames)
//          <dir>/samples.bin   F x C x (S * P) complex64 (frame, coil, sample)
//   writes traj.bin  [S * P, F] float2      golden-ratio trajectory
//          dcw.bin   [S * P]    float       golden-ratio radial weights (alpha 1.5)
//          sdc.bin   [S * P]    float       estimate_dcw on frame 0 (os 2.1, 10 its)
//          flags.txt            cycle_completed of each frame
//          acc.bin   [N, N, C]  complex64   accumulated coil images
//          csm.bin   [N, N, C]  complex64   estimate_b1_map of acc
//          comb.bin  [N, N]     complex64   combined coil image
//          mh.bin    [N, N]     complex64   E^H of frame 0 (sqrt dcw)
//          cg.bin    [N, N]     complex64   5 CG iterations on frame 0
//          cgreg.bin [N, N]     complex64   5 CG iterations on frame 0 in the
//                               order of gpuCgSenseGadget::process
//                               (gadgets/pmri/gpuCgSenseGadget.cpp:158-190):
//                               cuImageOperator from the combined image,
//                               weight kappa 0.05, its image scaled by kappa in
//                               place, preconditioner 1/sqrt(sum |csm|^2 + that)
// and checks that invalid calls throw std::runtime_error.  Exit status 0 on success.
#include <cmath>
#include <complex>
#include <fstream>
#include <iostream>
#include <stdexcept>
#include <string>
#include <vector>

#include "gt_pmri.hpp"

namespace {

using cplx = std::complex<float>;
using float2v = std::array<float, 2>;

// A non-owning view (get_data_ptr / get_dimensions), as a caller would make
// one frame of a larger array.
template <class T>
struct View {
    T* p;
    std::vector<size_t> d;
    T* get_data_ptr() const { return p; }
    const std::vector<size_t>& get_dimensions() const { return d; }
    size_t get_number_of_elements() const {
        size_t n = 1;
        for (size_t v : d) n *= v;
        return n;
    }
};

template <class T>
std::vector<T> read_file(const std::string& path, size_t n) {
    std::vector<T> v(n);
    std::ifstream f(path, std::ios::binary);
    if (!f.read(reinterpret_cast<char*>(v.data()), std::streamsize(n * sizeof(T))))
        throw std::runtime_error("cannot read " + path);
    return v;
}

template <class T>
void write_file(const std::string& path, const std::vector<T>& v) {
    std::ofstream f(path, std::ios::binary);
    f.write(reinterpret_cast<const char*>(v.data()), std::streamsize(v.size() * sizeof(T)));
    if (!f) throw std::runtime_error("cannot write " + path);
}

template <class A>
void dump(const std::string& path, const A& a) {
    using T = typename std::remove_pointer<decltype(a.get_data_ptr())>::type;
    std::vector<T> h(a.get_number_of_elements());
    gt_amd::check(gt_memcpy_dtoh(h.data(), a.get_data_ptr(), h.size() * sizeof(T), nullptr));
    write_file(path, h);
}

template <class F>
bool throws(F&& f) {
    try {
        f();
    } catch (const std::runtime_error&) {
        return true;
    }
    return false;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc != 2) {
        std::cerr << "usage: pmri_roundtrip <dir>\n";
        return 2;
    }
    const std::string dir = argv[1];
    try {
        size_t N = 0, C = 0, S = 0, P = 0, F = 0;
        {
            std::ifstream m(dir + "/meta.txt");
            if (!(m >> N >> C >> S >> P >> F)) throw std::runtime_error("bad meta.txt");
        }
        const size_t M = S * P;
        using namespace gt_amd;

        // trajectory and weights (main_cg.cpp:110-130)
        auto traj = compute_radial_trajectory_golden_ratio_2d<float>(unsigned(S), unsigned(P), unsigned(F));
        auto dcw = compute_radial_dcw_golden_ratio_2d<float>(unsigned(S), unsigned(P), 1.5f, 1.0f / 1.5f);
        dump(dir + "/traj.bin", *traj);
        dump(dir + "/dcw.bin", *dcw);
        View<float2v> traj0{traj->get_data_ptr(), {M}};
        auto sdc = estimate_dcw<float, 2>(traj0, std::array<size_t, 2>{N, N}, 2.1f, 10, 5.5f);
        dump(dir + "/sdc.bin", *sdc);

        // frames into the buffer (main_cg.cpp:132-160)
        auto samples_h = read_file<cplx>(dir + "/samples.bin", F * C * M);
        cuNDArray<cplx> samples(std::vector<size_t>{M, C, F});
        samples.upload(samples_h.data());
        cuSenseBuffer<float, 2> buf;
        buf.setup({N, N}, {N * 3 / 2, N * 3 / 2}, 5.5f, unsigned(C), 2, 1);
        buf.set_dcw(dcw);
        std::ofstream flags(dir + "/flags.txt");
        for (size_t f = 0; f < F; f++) {
            View<cplx> sf{samples.get_data_ptr() + f * M * C, {M, C}};
            View<float2v> tf{traj->get_data_ptr() + f * M, {M}};
            flags << (buf.add_frame_data(&sf, &tf) ? 1 : 0) << "\n";
        }
        flags.close();
        auto acc = buf.get_accumulated_coil_images();
        dump(dir + "/acc.bin", *acc);
        std::cout << "normalization " << buf.get_normalization_factor() << "\n";

        // coil maps and the combined image (main_cg.cpp:162-170)
        auto csm = std::make_shared<cuNDArray<cplx>>(estimate_b1_map<float, 2>(*acc));
        dump(dir + "/csm.bin", *csm);
        buf.set_csm(csm);
        dump(dir + "/comb.bin", *buf.get_combined_coil_image());

        // SENSE operator and CG on frame 0 (main_cg.cpp:172-200)
        std::vector<float> w(M);
        gt_amd::check(gt_memcpy_dtoh(w.data(), dcw->get_data_ptr(), M * sizeof(float), nullptr));
        for (float& v : w) v = std::sqrt(v);
        auto sqrt_dcw = std::make_shared<cuNDArray<float>>(std::vector<size_t>{M});
        sqrt_dcw->upload(w.data());
        auto E = std::make_shared<cuNonCartesianSenseOperator<float, 2>>();
        E->setup({N, N}, {N * 3 / 2, N * 3 / 2}, 5.5f);
        View<float2v> t0{traj->get_data_ptr(), {M, 1}};
        E->preprocess(&t0);
        E->set_dcw(sqrt_dcw);
        E->set_csm(csm);
        View<cplx> d0{samples.get_data_ptr(), {M, 1, C}};
        cuNDArray<cplx> mh(std::vector<size_t>{N, N, 1});
        E->mult_MH(&d0, &mh);
        dump(dir + "/mh.bin", mh);
        cuCgSolver<float, 2> cg;
        cg.set_encoding_operator(E);
        cg.set_max_iterations(5);
        cg.set_tc_tolerance(1e-12f);
        auto x = cg.solve(&d0);
        dump(dir + "/cg.bin", *x);
        std::cout << "rq";
        for (double r : cg.get_residual_history()) std::cout << " " << r;
        std::cout << "\n";

        // the CG-SENSE gadget's solve: regulariser from the combined image
        const float kappa = 0.05f;
        auto comb = buf.get_combined_coil_image();
        auto R = std::make_shared<cuImageOperator<float>>();
        R->set_weight(kappa);
        R->compute(comb.get());
        {
            // R_diag = R->get(); *R_diag *= kappa: the operator's own image
            auto Rd = R->get();
            std::vector<float> h(Rd->get_number_of_elements());
            Rd->download(h.data());
            for (float& v : h) v *= kappa;
            Rd->upload(h.data());
        }
        auto pw = cg_sense_precond_weights(csm.get(), R->get().get(), 1.0f);
        cuCgSolver<float, 2> cgr;
        cgr.set_encoding_operator(E);
        cgr.add_regularization_operator(R);
        cgr.set_preconditioner_weights(pw);
        cgr.set_max_iterations(5);
        cgr.set_tc_tolerance(1e-12f);
        auto xr = cgr.solve(&d0);
        dump(dir + "/cgreg.bin", *xr);
        std::cout << "offset " << R->get_offset() << "\n";

        // error paths
        bool ok = true;
        View<cplx> wrong{samples.get_data_ptr(), {M, C + 1}};
        ok &= throws([&] { buf.add_frame_data(&wrong, &traj0); });
        View<cplx> flat{acc->get_data_ptr(), {N * N * C}};
        ok &= throws([&] { estimate_b1_map<float, 2>(flat); });
        cuCgSolver<float, 2> empty;
        ok &= throws([&] { empty.solve(&d0); });
        cuSenseBuffer<float, 2> nocsm;
        nocsm.setup({N, N}, {N * 3 / 2, N * 3 / 2}, 5.5f, unsigned(C), 1, 1);
        ok &= throws([&] { nocsm.get_combined_coil_image(); });
        if (!ok) {
            std::cerr << "an invalid call did not throw\n";
            return 1;
        }
        std::cout << "pmri_roundtrip: ok\n";
    } catch (const std::exception& e) {
        std::cerr << "pmri_roundtrip: " << e.what() << "\n";
        return 1;
    }
    return 0;
}
