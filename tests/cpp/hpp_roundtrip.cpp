// hpp_roundtrip.cpp -- drives libgtnufft.so through the C++ front end
// (include/gt_nufft.hpp), the way a Gadgetron C++ caller such as
// cuNonCartesianSenseOperator (mri/pmri/gpu/cuNonCartesianSenseOperator.cu:14-84)
// drives cuNFFT_plan<float, 2>.
//
// usage: hpp_roundtrip <dir>
//   reads  <dir>/meta.txt      "N M F C"
//          <dir>/traj.bin      M*F float2 in [-1/2, 1/2]
//          <dir>/samples.bin   M*F*C complex64
//          <dir>/dcw.bin       M*F float
//   writes <dir>/adjoint.bin   BACKWARDS_NC2C of samples, N*N*F*C complex64
//          <dir>/forward.bin   FORWARDS_C2NC of adjoint.bin, M*F*C complex64
//          <dir>/mhm.bin       mult_MH_M of adjoint.bin, N*N*F*C complex64
//          <dir>/op_mh2.bin    NFFTOperator: mult_MH of samples, then again with
//                              accumulate (twice adjoint.bin)
//          <dir>/op_m.bin      NFFTOperator: mult_M of adjoint.bin (= forward.bin)
//          <dir>/op_mhm2.bin   NFFTOperator: mult_MH_M of adjoint.bin, then again
//                              with accumulate (twice mhm.bin)
// and checks that invalid calls throw std::runtime_error.  Exit status 0 on success.
#include <complex>
#include <cstdio>
#include <memory>
#include <fstream>
#include <iostream>
#include <stdexcept>
#include <string>
#include <vector>

#include "gt_nufft.hpp"

namespace {

using cplx = std::complex<float>;

// A Gadgetron-style device array: get_dimensions() returns by value, as
// NDArray<T>::get_dimensions() does (core/NDArray.h:79).
template <class T>
struct ValueDimsArray {
    gt_amd::cuNDArray<T>* a;
    std::vector<size_t> get_dimensions() const { return a->get_dimensions(); }
    T* get_data_ptr() const { return a->get_data_ptr(); }
    size_t get_number_of_elements() const { return a->get_number_of_elements(); }
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

}  // namespace

int main(int argc, char** argv) {
    if (argc != 2) {
        std::cerr << "usage: hpp_roundtrip <dir>\n";
        return 2;
    }
    const std::string dir = argv[1];
    try {
        size_t N = 0, M = 0, F = 0, C = 0;
        {
            std::ifstream meta(dir + "/meta.txt");
            if (!(meta >> N >> M >> F >> C)) throw std::runtime_error("bad meta.txt");
        }
        using namespace gt_amd;
        auto plan = NFFT<cuNDArray, float, 2>::make_plan({N, N}, {N * 3 / 2, N * 3 / 2}, 5.5f);
        if (plan->get_matrix_size_os()[0] != N * 3 / 2 || plan->get_W() != 5.5f)
            throw std::runtime_error("getters disagree with the plan");

        cuNDArray<float> traj({2 * M, F});  // M*F float2 (x, y)
        traj.upload(read_file<float>(dir + "/traj.bin", 2 * M * F).data());
        // the trajectory as the reference passes it: an [M, F] array of vector_td<float, 2>
        struct TrajView {
            const cuNDArray<float>* data;
            std::vector<size_t> dims;
            std::vector<size_t> get_dimensions() const { return dims; }
            float* get_data_ptr() const { return data->get_data_ptr(); }
        } tv{&traj, {M, F}};
        plan->preprocess(tv, NFFT_prep_mode::ALL);

        cuNDArray<cplx> samples({M, F, C});
        samples.upload(read_file<cplx>(dir + "/samples.bin", M * F * C).data());
        cuNDArray<float> dcw({M, F});
        dcw.upload(read_file<float>(dir + "/dcw.bin", M * F).data());

        cuNDArray<cplx> image({N, N, F, C});
        plan->compute(samples, image, &dcw, NFFT_comp_mode::BACKWARDS_NC2C);
        std::vector<cplx> h_img(N * N * F * C);
        image.download(h_img.data());
        write_file(dir + "/adjoint.bin", h_img);

        cuNDArray<cplx> fwd({M, F, C});
        ValueDimsArray<cplx> img_v{&image}, fwd_v{&fwd};
        plan->compute(img_v, fwd_v, &dcw, NFFT_comp_mode::FORWARDS_C2NC);
        std::vector<cplx> h_fwd(M * F * C);
        fwd.download(h_fwd.data());
        write_file(dir + "/forward.bin", h_fwd);

        cuNDArray<cplx> mhm({N, N, F, C});
        plan->mult_MH_M(image, mhm, &dcw);
        std::vector<cplx> h_mhm(N * N * F * C);
        mhm.download(h_mhm.data());
        write_file(dir + "/mhm.bin", h_mhm);

        // NFFTOperator<cuNDArray, float, 2> over the same trajectory and weights
        NFFTOperator<float, 2> op;
        op.setup({N, N}, {N * 3 / 2, N * 3 / 2}, 5.5f);
        op.preprocess(tv);
        auto op_dcw = std::make_shared<cuNDArray<float>>(std::vector<size_t>{M, F});
        op_dcw->upload(read_file<float>(dir + "/dcw.bin", M * F).data());
        op.set_dcw(op_dcw);
        auto save = [&](const std::string& name, const cuNDArray<cplx>& a) {
            std::vector<cplx> h(a.get_number_of_elements());
            a.download(h.data());
            write_file(dir + "/" + name, h);
        };
        cuNDArray<cplx> op_img({N, N, F, C});
        op.mult_MH(&samples, &op_img);
        op.mult_MH(&samples, &op_img, true);
        save("op_mh2.bin", op_img);
        cuNDArray<cplx> op_fwd({M, F, C});
        op.mult_M(&image, &op_fwd);
        save("op_m.bin", op_fwd);
        cuNDArray<cplx> op_mhm({N, N, F, C});
        bool threw = false;
        try {
            op.mult_MH_M(&image, &op_mhm);  // codomain not set yet
        } catch (const std::runtime_error&) {
            threw = true;
        }
        if (!threw) throw std::runtime_error("NFFTOperator::mult_MH_M without a codomain did not throw");
        op.set_codomain_dimensions({M, F, C});
        op.mult_MH_M(&image, &op_mhm);
        op.mult_MH_M(&image, &op_mhm, true);
        save("op_mhm2.bin", op_mhm);
        threw = false;
        try {
            op.mult_M(static_cast<cuNDArray<cplx>*>(nullptr), &op_fwd);
        } catch (const std::runtime_error&) {
            threw = true;
        }
        if (!threw) throw std::runtime_error("NFFTOperator with a null input did not throw");

        // the reference's validation paths throw std::runtime_error
        threw = false;
        try {
            cuNDArray<cplx> wrong({N + 1, N, F, C});
            plan->compute(samples, wrong, &dcw, NFFT_comp_mode::BACKWARDS_NC2C);
        } catch (const std::runtime_error&) {
            threw = true;
        }
        if (!threw) throw std::runtime_error("a wrong image size did not throw");
        threw = false;
        try {
            auto bad = NFFT<cuNDArray, float, 2>::make_plan({N, N}, {N * 3 / 2, N * 3 / 2}, 1.0f);
        } catch (const std::runtime_error&) {
            threw = true;
        }
        if (!threw) throw std::runtime_error("a kernel width below 1.8 did not throw");
        std::cout << "hpp_roundtrip: ok\n";
        return 0;
    } catch (const std::exception& e) {
        std::cerr << "hpp_roundtrip: " << e.what() << "\n";
        return 1;
    }
}
