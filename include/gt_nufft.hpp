// gt_nufft.hpp -- header-only C++ front end of libgtnufft.so with the shape of
// Gadgetron's cuNFFT_plan<REAL,D> API (toolboxes/nfft/NFFT.h:54-175,
// toolboxes/nfft/gpu/cuNFFT.h:52-131), so that C++ callers written against the
// reference (a non-Cartesian SENSE operator, a CG solver) switch by changing
// includes and the namespace.
//
//   using namespace gt_amd;
//   auto plan = NFFT<cuNDArray, float, 2>::make_plan(N, N_os, W, ConvolutionType::STANDARD);
//   plan->preprocess(traj, NFFT_prep_mode::ALL);
//   plan->compute(in, out, &dcw, NFFT_comp_mode::BACKWARDS_NC2C);
//
// The ARRAY template parameter of every operation is duck typed: anything with
// get_data_ptr() and get_dimensions() (returning x-fastest dims) works,
// including Gadgetron's own cuNDArray<T>.  gt_amd::cuNDArray<T> below is a
// minimal owning device array for callers without one.  Failing calls throw
// std::runtime_error carrying the library's message, as the reference does.
#pragma once

#include <array>
#include <complex>
#include <memory>
#include <numeric>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "gt_nufft.h"

namespace gt_amd {

enum class NFFT_fft_mode { FORWARDS = GT_FFT_FORWARDS, BACKWARDS = GT_FFT_BACKWARDS };
enum class NFFT_conv_mode { C2NC = GT_CONV_C2NC, NC2C = GT_CONV_NC2C };
enum class NFFT_comp_mode {
    FORWARDS_C2NC = GT_FORWARDS_C2NC,
    FORWARDS_NC2C = GT_FORWARDS_NC2C,
    BACKWARDS_C2NC = GT_BACKWARDS_C2NC,
    BACKWARDS_NC2C = GT_BACKWARDS_NC2C
};
enum class NFFT_prep_mode { C2NC = GT_PREP_C2NC, NC2C = GT_PREP_NC2C, ALL = GT_PREP_ALL };
enum class ConvolutionType { STANDARD = GT_CONV_STANDARD, ATOMIC = GT_CONV_ATOMIC, SPARSE_MATRIX = GT_CONV_SPARSE_MATRIX };

inline void check(gt_status s) {
    if (s != GT_OK) throw std::runtime_error(std::string("gt_nufft: ") + gt_last_error());
}

// Minimal owning device array (x-fastest dims), the role of cuNDArray<T>.
template <class T>
class cuNDArray {
public:
    using element_type = T;
    cuNDArray() = default;
    explicit cuNDArray(std::vector<size_t> dims) : dims_(std::move(dims)) { allocate(); }
    cuNDArray(const cuNDArray&) = delete;
    cuNDArray& operator=(const cuNDArray&) = delete;
    cuNDArray(cuNDArray&& o) noexcept : dims_(std::move(o.dims_)), ptr_(o.ptr_) { o.ptr_ = nullptr; }
    ~cuNDArray() { if (ptr_) gt_free(ptr_); }

    T* get_data_ptr() const { return ptr_; }
    const std::vector<size_t>& get_dimensions() const { return dims_; }
    size_t get_number_of_dimensions() const { return dims_.size(); }
    size_t get_size(size_t d) const { return d < dims_.size() ? dims_[d] : 1; }
    size_t get_number_of_elements() const {
        return std::accumulate(dims_.begin(), dims_.end(), size_t(1), std::multiplies<size_t>());
    }
    void upload(const T* host) { check(gt_memcpy_htod(ptr_, host, bytes(), nullptr)); }
    void download(T* host) const { check(gt_memcpy_dtoh(host, ptr_, bytes(), nullptr)); }
    void clear() {
        check(gt_memset(ptr_, 0, bytes(), nullptr));
        check(gt_stream_synchronize(nullptr));
    }

private:
    size_t bytes() const { return get_number_of_elements() * sizeof(T); }
    void allocate() {
        void* p = nullptr;
        check(gt_malloc(&p, std::max<size_t>(bytes(), 1)));
        ptr_ = static_cast<T*>(p);
    }
    std::vector<size_t> dims_;
    T* ptr_ = nullptr;
};

namespace detail {
template <class A>
std::vector<size_t> dims_of(const A& a) {
    const auto& d = a.get_dimensions();
    return std::vector<size_t>(d.begin(), d.end());
}
}  // namespace detail

// cuNFFT_plan<REAL, D>: REAL = float, or double (STANDARD convolution only; the
// DCW array of a double plan holds doubles).
template <class REAL, unsigned int D>
class cuNFFT_plan {
    static_assert(D >= 1 && D <= 4, "D must be 1 to 4");

public:
    using vec = std::array<size_t, D>;

    // NFFT_plan(N, N_os, W) (NFFT.hpp:6-20)
    cuNFFT_plan(const vec& matrix_size, const vec& matrix_size_os, REAL W,
                ConvolutionType conv = ConvolutionType::STANDARD, int device = -1) {
        gt_nufft_desc d{};
        fill(d, matrix_size, W, conv, device);
        for (unsigned i = 0; i < D; i++) d.N_os[i] = matrix_size_os[i];
        create(d);
    }
    // NFFT_plan(N, os_factor, W) (NFFT.hpp:23-37)
    cuNFFT_plan(const vec& matrix_size, REAL os_factor, REAL W,
                ConvolutionType conv = ConvolutionType::STANDARD, int device = -1) {
        gt_nufft_desc d{};
        fill(d, matrix_size, W, conv, device);
        d.os_factor = os_factor;
        create(d);
    }
    cuNFFT_plan(const cuNFFT_plan&) = delete;
    cuNFFT_plan& operator=(const cuNFFT_plan&) = delete;
    virtual ~cuNFFT_plan() { if (h_) gt_nufft_destroy(h_); }

    template <class TRAJ>
    void preprocess(const TRAJ& trajectory, NFFT_prep_mode mode = NFFT_prep_mode::ALL,
                    gt_stream_t stream = nullptr) {
        auto d = detail::dims_of(trajectory);
        const size_t M = d.empty() ? 0 : d[0];
        size_t F = 1;
        for (size_t i = 1; i < d.size(); i++) F *= d[i];
        check(gt_nufft_preprocess(h_, trajectory.get_data_ptr(), M, F, int(mode), stream));
    }

    template <class IN, class OUT, class DCW>
    void compute(const IN& in, OUT& out, const DCW* dcw, NFFT_comp_mode mode, gt_stream_t stream = nullptr) {
        auto di = detail::dims_of(in), dout = detail::dims_of(out);
        const float* w = dcw ? reinterpret_cast<const float*>(dcw->get_data_ptr()) : nullptr;
        const size_t wl = dcw ? dcw->get_number_of_elements() : 0;
        check(gt_nufft_compute(h_, in.get_data_ptr(), di.data(), int(di.size()), out.get_data_ptr(),
                               dout.data(), int(dout.size()), w, wl, int(mode), stream));
    }
    template <class IN, class OUT>
    void compute(const IN& in, OUT& out, std::nullptr_t, NFFT_comp_mode mode, gt_stream_t stream = nullptr) {
        auto di = detail::dims_of(in), dout = detail::dims_of(out);
        check(gt_nufft_compute(h_, in.get_data_ptr(), di.data(), int(di.size()), out.get_data_ptr(),
                               dout.data(), int(dout.size()), nullptr, 0, int(mode), stream));
    }

    template <class IN, class OUT, class DCW>
    void mult_MH_M(const IN& in, OUT& out, const DCW* dcw, gt_stream_t stream = nullptr) {
        auto d = detail::dims_of(in);
        const float* w = dcw ? reinterpret_cast<const float*>(dcw->get_data_ptr()) : nullptr;
        const size_t wl = dcw ? dcw->get_number_of_elements() : 0;
        check(gt_nufft_mult_MH_M(h_, in.get_data_ptr(), out.get_data_ptr(), d.data(), int(d.size()), w, wl, stream));
    }

    template <class IN, class OUT>
    void convolve(const IN& in, OUT& out, NFFT_conv_mode mode, bool accumulate = false, gt_stream_t stream = nullptr) {
        auto di = detail::dims_of(in), dout = detail::dims_of(out);
        check(gt_nufft_convolve(h_, in.get_data_ptr(), di.data(), int(di.size()), out.get_data_ptr(),
                                dout.data(), int(dout.size()), int(mode), accumulate ? 1 : 0, stream));
    }

    template <class A>
    void fft(A& data, NFFT_fft_mode mode, bool do_scale = true, gt_stream_t stream = nullptr) {
        auto d = detail::dims_of(data);
        check(gt_nufft_fft(h_, data.get_data_ptr(), d.data(), int(d.size()), int(mode), do_scale ? 1 : 0, stream));
    }

    template <class A>
    void deapodize(A& image, bool fourier_domain = false, gt_stream_t stream = nullptr) {
        auto d = detail::dims_of(image);
        check(gt_nufft_deapodize(h_, image.get_data_ptr(), d.data(), int(d.size()), fourier_domain ? 1 : 0, stream));
    }

    vec get_matrix_size() const {
        size_t n[4] = {0, 0, 0, 0};
        check(gt_nufft_get_matrix_size(h_, n));
        vec v{};
        for (unsigned i = 0; i < D; i++) v[i] = n[i];
        return v;
    }
    vec get_matrix_size_os() const {
        size_t n[4] = {0, 0, 0, 0};
        check(gt_nufft_get_matrix_size_os(h_, n));
        vec v{};
        for (unsigned i = 0; i < D; i++) v[i] = n[i];
        return v;
    }
    REAL get_W() const {
        double w = 0;
        check(gt_nufft_get_W(h_, &w));
        return REAL(w);
    }
    gt_nufft_plan* handle() const { return h_; }

private:
    static void fill(gt_nufft_desc& d, const vec& N, REAL W, ConvolutionType conv, int device) {
        d.D = int(D);
        d.precision = std::is_same<REAL, double>::value ? GT_DOUBLE : GT_FLOAT;
        for (unsigned i = 0; i < D; i++) d.N[i] = N[i];
        d.W = W;
        d.conv_type = int(conv);
        d.device = device;
    }
    void create(const gt_nufft_desc& d) { check(gt_nufft_create(&h_, &d)); }
    gt_nufft_plan* h_ = nullptr;
};

// NFFT<cuNDArray, REAL, D>::make_plan (cuNFFT.h:113-131), returning a shared plan.
template <template <class> class ARRAY, class REAL, unsigned int D>
struct NFFT {
    static std::shared_ptr<cuNFFT_plan<REAL, D>> make_plan(
        const std::array<size_t, D>& matrix_size, const std::array<size_t, D>& matrix_size_os, REAL W,
        ConvolutionType conv = ConvolutionType::STANDARD) {
        return std::make_shared<cuNFFT_plan<REAL, D>>(matrix_size, matrix_size_os, W, conv);
    }
};

// NFFTOperator<cuNDArray, REAL, D> (toolboxes/nfft/NFFTOperator.h:10-37,
// NFFTOperator.hpp:13-98): a linear operator over one plan and an optional
// dcw.  mult_M = compute(FORWARDS_C2NC), mult_MH = compute(BACKWARDS_NC2C),
// mult_MH_M = the plan's mult_MH_M, each with the operator's dcw; accumulate
// computes into a temporary and adds it to out (gt_array_add), as the
// reference does with operator+=.
template <class REAL, unsigned int D>
class NFFTOperator {
public:
    using vec = std::array<size_t, D>;
    using plan_type = cuNFFT_plan<REAL, D>;

    virtual ~NFFTOperator() = default;

    // setup (NFFTOperator.hpp:84-89): make_plan(N, N_os, W)
    virtual void setup(const vec& matrix_size, const vec& matrix_size_os, REAL W) {
        plan_ = NFFT<cuNDArray, REAL, D>::make_plan(matrix_size, matrix_size_os, W);
    }
    // preprocess (NFFTOperator.hpp:91-95): NFFT_prep_mode::ALL
    template <class TRAJ>
    void preprocess(const TRAJ& trajectory) {
        need_plan("preprocess");
        plan_->preprocess(trajectory, NFFT_prep_mode::ALL);
    }
    virtual void set_dcw(std::shared_ptr<cuNDArray<REAL>> dcw) { dcw_ = std::move(dcw); }
    std::shared_ptr<cuNDArray<REAL>> get_dcw() { return dcw_; }
    std::shared_ptr<plan_type> get_plan() { return plan_; }

    // linearOperator's domain / codomain (operators/linearOperator.h)
    void set_domain_dimensions(std::vector<size_t> d) { domain_ = std::move(d); }
    void set_codomain_dimensions(std::vector<size_t> d) { codomain_ = std::move(d); }
    std::vector<size_t> get_domain_dimensions() const { return domain_; }
    std::vector<size_t> get_codomain_dimensions() const { return codomain_; }

    template <class IN, class OUT>
    void mult_M(IN* in, OUT* out, bool accumulate = false) {
        apply("mult_M", in, out, accumulate, [&](auto& dst) {
            plan_->compute(*in, dst, dcw_.get(), NFFT_comp_mode::FORWARDS_C2NC);
        });
    }
    template <class IN, class OUT>
    void mult_MH(IN* in, OUT* out, bool accumulate = false) {
        apply("mult_MH", in, out, accumulate, [&](auto& dst) {
            plan_->compute(*in, dst, dcw_.get(), NFFT_comp_mode::BACKWARDS_NC2C);
        });
    }
    // NFFTOperator.hpp:59-80: refuses to run before the codomain is set
    template <class IN, class OUT>
    void mult_MH_M(IN* in, OUT* out, bool accumulate = false) {
        if (in && out && codomain_.empty())
            throw std::runtime_error("NFFTOperator::mult_MH_M : operator codomain dimensions not set");
        apply("mult_MH_M", in, out, accumulate, [&](auto& dst) { plan_->mult_MH_M(*in, dst, dcw_.get()); });
    }

protected:
    void need_plan(const char* who) const {
        if (!plan_) throw std::runtime_error(std::string("NFFTOperator::") + who + " : setup has not been called");
    }
    template <class IN, class OUT, class F>
    void apply(const char* who, IN* in, OUT* out, bool accumulate, F&& run) {
        if (!in || !out)
            throw std::runtime_error(std::string("NFFTOperator::") + who + " : 0x0 input/output not accepted");
        need_plan(who);
        if (!accumulate) {
            run(*out);
            return;
        }
        cuNDArray<std::complex<REAL>> tmp(detail::dims_of(*out));
        run(tmp);
        check(gt_array_add(out->get_data_ptr(), tmp.get_data_ptr(), 2 * tmp.get_number_of_elements(),
                           std::is_same<REAL, double>::value ? GT_DOUBLE : GT_FLOAT, nullptr));
        check(gt_stream_synchronize(nullptr));  // tmp is freed on return
    }

    std::shared_ptr<plan_type> plan_;
    std::shared_ptr<cuNDArray<REAL>> dcw_;
    std::vector<size_t> domain_, codomain_;
};

}  // namespace gt_amd
