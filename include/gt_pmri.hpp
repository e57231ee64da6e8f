// gt_pmri.hpp -- header-only C++ front end of the calls around the NFFT on the
// non-Cartesian SENSE path, in the shape of the reference's APIs, so that C++
// callers written against it (the radial SENSE gadgets, the standalone CG
// applications) switch by changing includes and the namespace:
//
//   core/gpu/radial_utilities.h:13-38         trajectory and radial DCW generators
//   mri/sdc/gpu/cuSDC.h:20-44                 estimate_dcw
//   mri/pmri/gpu/cuBuffer.h:13-57             cuBuffer<REAL, D>
//   mri/pmri/gpu/cuSenseBuffer.h:8-34         cuSenseBuffer<REAL, D>
//   mri/pmri/gpu/b1_map.h:16-24               estimate_b1_map<REAL, D>
//   mri/pmri/gpu/cuNonCartesianSenseOperator.h:12-44, cuSenseOperator.h
//                                             cuNonCartesianSenseOperator<REAL, D>
//   solvers/gpu/cuCgSolver.h, solvers/cgSolver.h:105-390
//                                             cuCgSolver<T> for that operator
//   operators/gpu/cuImageOperator.h:16-66     cuImageOperator<T> (image regulariser)
//
// Arrays are duck typed as in gt_nufft.hpp (get_data_ptr(), get_dimensions()
// with x fastest, get_number_of_elements()); results come back as
// std::shared_ptr<gt_amd::cuNDArray<...>> where the reference returns
// boost::shared_ptr<cuNDArray<...>>.  REAL = float.  Failing calls throw
// std::runtime_error with the library's message.
#pragma once

#include <array>
#include <complex>
#include <memory>
#include <stdexcept>
#include <vector>

#include "gt_b1map.h"
#include "gt_buffer.h"
#include "gt_cgsense.h"
#include "gt_nufft.hpp"
#include "gt_radial.h"

namespace gt_amd {

// GOLDEN_RATIO_ANGULAR_STEP_SIZE (radial_utilities.h:10-13)
enum GOLDEN_RATIO_ANGULAR_STEP_SIZE { GR_SMALLEST = GT_GR_SMALLEST, GR_ORIGINAL = GT_GR_ORIGINAL };

namespace detail {
template <class REAL>
void require_float() {
    static_assert(std::is_same<REAL, float>::value, "libgtnufft computes in float");
}
template <class A>
size_t elements(const A& a) {
    size_t n = 1;
    for (size_t d : a.get_dimensions()) n *= d;
    return n;
}
}  // namespace detail

// ---------------------------------------------------------------------------
// Trajectories and density compensation
// ---------------------------------------------------------------------------

// compute_radial_trajectory_golden_ratio_2d: [S * P, F] of REAL 2-vectors.
template <class REAL>
std::shared_ptr<cuNDArray<std::array<REAL, 2>>> compute_radial_trajectory_golden_ratio_2d(
    unsigned int num_samples_per_profile, unsigned int num_profiles_per_frame, unsigned int num_frames,
    unsigned int profile_offset = 0, GOLDEN_RATIO_ANGULAR_STEP_SIZE mode = GR_ORIGINAL) {
    detail::require_float<REAL>();
    auto out = std::make_shared<cuNDArray<std::array<REAL, 2>>>(
        std::vector<size_t>{size_t(num_samples_per_profile) * num_profiles_per_frame, num_frames});
    check(gt_radial_trajectory_golden_ratio_2d(reinterpret_cast<float*>(out->get_data_ptr()), num_samples_per_profile,
                                               num_profiles_per_frame, num_frames, profile_offset, int(mode),
                                               nullptr));
    return out;
}

// compute_radial_trajectory_fixed_angle_2d: [S * P, F].
template <class REAL>
std::shared_ptr<cuNDArray<std::array<REAL, 2>>> compute_radial_trajectory_fixed_angle_2d(
    unsigned int num_samples_per_profile, unsigned int num_profiles_per_frame, unsigned int num_frames,
    REAL angular_offset = REAL(0)) {
    detail::require_float<REAL>();
    auto out = std::make_shared<cuNDArray<std::array<REAL, 2>>>(
        std::vector<size_t>{size_t(num_samples_per_profile) * num_profiles_per_frame, num_frames});
    check(gt_radial_trajectory_fixed_angle_2d(reinterpret_cast<float*>(out->get_data_ptr()), num_samples_per_profile,
                                              num_profiles_per_frame, num_frames, angular_offset, nullptr));
    return out;
}

// compute_radial_trajectory_variable_angle_2d: one angle per frame (the
// reference kernel's indexing); angular_offset is unused there as here.
template <class REAL, class A>
std::shared_ptr<cuNDArray<std::array<REAL, 2>>> compute_radial_trajectory_variable_angle_2d(
    A* angles, unsigned int num_samples_per_profile, unsigned int num_profiles_per_frame, unsigned int num_frames,
    REAL angular_offset = REAL(0)) {
    detail::require_float<REAL>();
    (void)angular_offset;
    if (!angles || detail::elements(*angles) < num_frames)
        throw std::runtime_error("compute_radial_trajectory_variable_angle_2d: one angle per frame required");
    auto out = std::make_shared<cuNDArray<std::array<REAL, 2>>>(
        std::vector<size_t>{size_t(num_samples_per_profile) * num_profiles_per_frame, num_frames});
    check(gt_radial_trajectory_variable_angle_2d(reinterpret_cast<float*>(out->get_data_ptr()),
                                                 reinterpret_cast<const float*>(angles->get_data_ptr()),
                                                 num_samples_per_profile, num_profiles_per_frame, num_frames,
                                                 nullptr));
    return out;
}

// compute_radial_dcw_golden_ratio_2d: [S * P].
template <class REAL>
std::shared_ptr<cuNDArray<REAL>> compute_radial_dcw_golden_ratio_2d(
    unsigned int num_samples_per_profile, unsigned int num_profiles, REAL alpha,
    REAL one_over_radial_oversampling_factor, unsigned int profile_offset = 0,
    GOLDEN_RATIO_ANGULAR_STEP_SIZE mode = GR_ORIGINAL) {
    detail::require_float<REAL>();
    auto out = std::make_shared<cuNDArray<REAL>>(std::vector<size_t>{size_t(num_samples_per_profile) * num_profiles});
    check(gt_radial_dcw_golden_ratio_2d(out->get_data_ptr(), num_samples_per_profile, num_profiles, alpha,
                                        one_over_radial_oversampling_factor, profile_offset, int(mode), nullptr));
    return out;
}

// compute_radial_dcw_fixed_angle_2d: [S * P].
template <class REAL>
std::shared_ptr<cuNDArray<REAL>> compute_radial_dcw_fixed_angle_2d(unsigned int num_samples_per_profile,
                                                                  unsigned int num_profiles, REAL alpha,
                                                                  REAL one_over_radial_oversampling_factor) {
    detail::require_float<REAL>();
    auto out = std::make_shared<cuNDArray<REAL>>(std::vector<size_t>{size_t(num_samples_per_profile) * num_profiles});
    check(gt_radial_dcw_fixed_angle_2d(out->get_data_ptr(), num_samples_per_profile, num_profiles, alpha,
                                       one_over_radial_oversampling_factor, nullptr));
    return out;
}

// estimate_dcw (cuSDC.h:20-44, the GPU defaults: os_factor 1.5, 10
// iterations, width 5.5): weights of the trajectory's shape.  The trajectory
// is one flattened frame of REAL D-vectors.
template <class REAL, unsigned int D, class TRAJ>
std::shared_ptr<cuNDArray<REAL>> estimate_dcw(const TRAJ& traj, const std::array<size_t, D>& matrix_size,
                                              REAL os_factor = REAL(1.5), unsigned int num_iterations = 10,
                                              REAL kernelWidth = REAL(5.5)) {
    detail::require_float<REAL>();
    auto out = std::make_shared<cuNDArray<REAL>>(std::vector<size_t>(traj.get_dimensions().begin(),
                                                                     traj.get_dimensions().end()));
    check(gt_sdc_estimate_dcw(reinterpret_cast<const float*>(traj.get_data_ptr()), detail::elements(traj), int(D),
                              matrix_size.data(), os_factor, num_iterations, kernelWidth, nullptr,
                              out->get_data_ptr(), nullptr));
    return out;
}

// estimate_dcw with initial weights (cuSDC.h:41-44).
template <class REAL, unsigned int D, class TRAJ, class W>
std::shared_ptr<cuNDArray<REAL>> estimate_dcw(const TRAJ& traj, const W& initial_dcw,
                                              const std::array<size_t, D>& matrix_size, REAL os_factor = REAL(1.5),
                                              unsigned int num_iterations = 10, REAL kernelWidth = REAL(5.5)) {
    detail::require_float<REAL>();
    if (detail::elements(initial_dcw) != detail::elements(traj))
        throw std::runtime_error("estimate_dcw: initial_dcw and trajectory differ in size");
    auto out = std::make_shared<cuNDArray<REAL>>(std::vector<size_t>(traj.get_dimensions().begin(),
                                                                     traj.get_dimensions().end()));
    check(gt_sdc_estimate_dcw(reinterpret_cast<const float*>(traj.get_data_ptr()), detail::elements(traj), int(D),
                              matrix_size.data(), os_factor, num_iterations, kernelWidth,
                              reinterpret_cast<const float*>(initial_dcw.get_data_ptr()), out->get_data_ptr(),
                              nullptr));
    return out;
}

// ---------------------------------------------------------------------------
// Coil sensitivities
// ---------------------------------------------------------------------------

// estimate_b1_map<REAL, D> (b1_map.h:16-24): data [N..., C] -> [N..., T].
template <class REAL, unsigned int D, class A>
cuNDArray<std::complex<REAL>> estimate_b1_map(const A& data, int target_coils = -1) {
    detail::require_float<REAL>();
    std::vector<size_t> dims(data.get_dimensions().begin(), data.get_dimensions().end());
    if (dims.size() != D + 1) throw std::runtime_error("estimate_b1_map:: dimensionality mismatch.");
    const size_t C = dims[D];
    const size_t T = (target_coils <= 0 || size_t(target_coils) > C) ? C : size_t(target_coils);
    std::vector<size_t> odims(dims);
    odims[D] = T;
    cuNDArray<std::complex<REAL>> out(odims);
    check(gt_estimate_b1_map(data.get_data_ptr(), int(D), dims.data(), unsigned(C), target_coils, out.get_data_ptr(),
                             nullptr));
    return out;
}

// ---------------------------------------------------------------------------
// Frame accumulation
// ---------------------------------------------------------------------------

// cuBuffer<REAL, D> (cuBuffer.h:13-57).  get_combined_coil_image is
// cuSenseBuffer's; here it is available once set_csm has been called.
template <class REAL, unsigned int D>
class cuBuffer {
    static_assert(D >= 1 && D <= 3, "D must be 1, 2 or 3");

public:
    using _uint64d = std::array<size_t, D>;
    cuBuffer() {
        detail::require_float<REAL>();
        check(gt_buffer_create(&h_));
    }
    cuBuffer(const cuBuffer&) = delete;
    cuBuffer& operator=(const cuBuffer&) = delete;
    virtual ~cuBuffer() {
        if (h_) gt_buffer_destroy(h_);
    }

    virtual void setup(_uint64d matrix_size, _uint64d matrix_size_os, REAL W, unsigned int num_coils,
                       unsigned int num_cycles, unsigned int num_sub_cycles) {
        check(gt_buffer_setup(h_, int(D), matrix_size.data(), matrix_size_os.data(), W, num_coils, num_cycles,
                              num_sub_cycles));
        matrix_size_ = matrix_size;
        num_coils_ = num_coils;
    }
    // Kept by reference, as the reference keeps the shared pointer.
    template <class A>
    void set_dcw(std::shared_ptr<A> dcw) {
        dcw_keep_ = dcw;
        check(gt_buffer_set_dcw(h_, dcw ? reinterpret_cast<const float*>(dcw->get_data_ptr()) : nullptr,
                                dcw ? detail::elements(*dcw) : 0));
    }
    virtual void clear() { check(gt_buffer_clear(h_, nullptr)); }

    // samples: [M, C] complex; trajectory: [M] REAL D-vectors.
    template <class S, class T>
    bool add_frame_data(S* samples, T* trajectory) {
        if (!samples || !trajectory) throw std::runtime_error("cuBuffer::add_frame_data: illegal input pointer");
        std::vector<size_t> sd(samples->get_dimensions().begin(), samples->get_dimensions().end());
        if (sd.empty() || sd.back() != num_coils_)
            throw std::runtime_error("cuBuffer::add_frame_data: unexpected number of coils according to setup");
        const size_t M = detail::elements(*samples) / num_coils_;
        if (detail::elements(*trajectory) != M)
            throw std::runtime_error("cuBuffer::add_frame_data: trajectory and samples disagree");
        int done = 0;
        check(gt_buffer_add_frame_data(h_, samples->get_data_ptr(), M,
                                       reinterpret_cast<const float*>(trajectory->get_data_ptr()), &done, nullptr));
        return done != 0;
    }

    virtual std::shared_ptr<cuNDArray<std::complex<REAL>>> get_accumulated_coil_images() {
        auto out = std::make_shared<cuNDArray<std::complex<REAL>>>(image_dims(num_coils_));
        check(gt_buffer_get_accumulated_coil_images(h_, out->get_data_ptr(), nullptr));
        return out;
    }
    REAL get_normalization_factor() const {
        float f = 0.0f;
        check(gt_buffer_get_normalization_factor(h_, &f));
        return f;
    }
    gt_buffer* handle() const { return h_; }

protected:
    std::vector<size_t> image_dims(size_t last) const {
        std::vector<size_t> d(matrix_size_.begin(), matrix_size_.end());
        d.push_back(last);
        return d;
    }
    gt_buffer* h_ = nullptr;
    _uint64d matrix_size_{};
    unsigned int num_coils_ = 0;
    std::shared_ptr<void> dcw_keep_;
};

// cuSenseBuffer<REAL, D> (cuSenseBuffer.h:8-34).
template <class REAL, unsigned int D>
class cuSenseBuffer : public cuBuffer<REAL, D> {
public:
    template <class A>
    void set_csm(std::shared_ptr<A> csm) {
        csm_keep_ = csm;
        csm_ = csm ? static_cast<const void*>(csm->get_data_ptr()) : nullptr;
    }
    // sum_c conj(csm_c) image_c of the accumulated coil images: [N...].
    virtual std::shared_ptr<cuNDArray<std::complex<REAL>>> get_combined_coil_image() {
        if (!csm_) throw std::runtime_error("cuSenseBuffer::get_combined_coil_image: csm not set");
        std::vector<size_t> d(this->matrix_size_.begin(), this->matrix_size_.end());
        auto out = std::make_shared<cuNDArray<std::complex<REAL>>>(d);
        check(gt_buffer_get_combined_coil_image(this->h_, csm_, out->get_data_ptr(), nullptr));
        return out;
    }

private:
    std::shared_ptr<void> csm_keep_;
    const void* csm_ = nullptr;
};

// ---------------------------------------------------------------------------
// SENSE operator and CG solver
// ---------------------------------------------------------------------------

// cuNonCartesianSenseOperator<REAL, D> (cuNonCartesianSenseOperator.h:12-44):
// owns its plan; set_dcw takes sqrt(DCW) as the reference's callers pass it.
template <class REAL, unsigned int D>
class cuNonCartesianSenseOperator {
    static_assert(D == 2 || D == 3, "the SENSE operator is 2D or 3D");

public:
    using _uint64d = std::array<size_t, D>;
    explicit cuNonCartesianSenseOperator(ConvolutionType conv = ConvolutionType::STANDARD) : conv_(conv) {}
    cuNonCartesianSenseOperator(const cuNonCartesianSenseOperator&) = delete;
    cuNonCartesianSenseOperator& operator=(const cuNonCartesianSenseOperator&) = delete;
    virtual ~cuNonCartesianSenseOperator() {
        if (op_) gt_sense_destroy(op_);
    }

    virtual void setup(_uint64d matrix_size, _uint64d matrix_size_os, REAL W) {
        plan_ = NFFT<cuNDArray, REAL, D>::make_plan(matrix_size, matrix_size_os, W, conv_);
        matrix_size_ = matrix_size;
        invalidate();
    }
    template <class TRAJ>
    void preprocess(TRAJ* trajectory) {
        if (!plan_) throw std::runtime_error("cuNonCartesianSenseOperator::preprocess: setup not called");
        if (!trajectory) throw std::runtime_error("cuNonCartesianSenseOperator::preprocess: illegal trajectory");
        plan_->preprocess(*trajectory, NFFT_prep_mode::ALL);
        const auto& d = trajectory->get_dimensions();
        frames_ = 1;
        for (size_t i = 1; i < d.size(); i++) frames_ *= d[i];
        samples_ = d.empty() ? 0 : d[0];
        invalidate();
    }
    template <class A>
    void set_dcw(std::shared_ptr<A> dcw) {
        dcw_keep_ = dcw;
        dcw_ = dcw ? reinterpret_cast<const float*>(dcw->get_data_ptr()) : nullptr;
        dcw_len_ = dcw ? detail::elements(*dcw) : 0;
        invalidate();
    }
    // cuSenseOperator::set_csm: [N..., C].
    template <class A>
    void set_csm(std::shared_ptr<A> csm) {
        csm_keep_ = csm;
        csm_ = csm ? static_cast<const void*>(csm->get_data_ptr()) : nullptr;
        coils_ = csm ? csm->get_dimensions().back() : 0;
        invalidate();
    }

    template <class IN, class OUT>
    void mult_M(IN* in, OUT* out, bool accumulate = false) {
        check(gt_sense_mult_M(op(), in->get_data_ptr(), out->get_data_ptr(), accumulate ? 1 : 0, nullptr));
    }
    template <class IN, class OUT>
    void mult_MH(IN* in, OUT* out, bool accumulate = false) {
        check(gt_sense_mult_MH(op(), in->get_data_ptr(), out->get_data_ptr(), accumulate ? 1 : 0, nullptr));
    }
    template <class IN, class OUT>
    void mult_MH_M(IN* in, OUT* out, bool accumulate = false) {
        check(gt_sense_mult_MH_M(op(), in->get_data_ptr(), out->get_data_ptr(), accumulate ? 1 : 0, nullptr));
    }
    // The image domain [N..., F].
    std::vector<size_t> get_domain_dimensions() const {
        std::vector<size_t> d(matrix_size_.begin(), matrix_size_.end());
        d.push_back(frames_);
        return d;
    }
    // The encoding operator weight (linearOperator::set_weight).
    void set_weight(REAL w) { weight_ = w; }
    REAL get_weight() const { return weight_; }
    gt_sense_op* op() {
        if (!op_) {
            if (!plan_ || !csm_) throw std::runtime_error("cuNonCartesianSenseOperator: setup, preprocess and set_csm first");
            check(gt_sense_create(&op_, plan_->handle(), csm_, coils_, dcw_, dcw_len_));
        }
        return op_;
    }

private:
    void invalidate() {
        if (op_) gt_sense_destroy(op_);
        op_ = nullptr;
    }
    ConvolutionType conv_;
    std::shared_ptr<cuNFFT_plan<REAL, D>> plan_;
    _uint64d matrix_size_{};
    size_t frames_ = 1, samples_ = 0, coils_ = 0;
    std::shared_ptr<void> dcw_keep_, csm_keep_;
    const float* dcw_ = nullptr;
    size_t dcw_len_ = 0;
    const void* csm_ = nullptr;
    REAL weight_ = REAL(1);
    gt_sense_op* op_ = nullptr;
};

// cuImageOperator<float_complext> (operators/imageOperator.h:14-96,
// operators/gpu/cuImageOperator.h:16-66): a regulariser with the real image
// 1 / (|x| n / asum(x) + offset) computed from a complex image x.  get()
// returns the operator's own image array, as the reference's shared_ptr does:
// scaling it in place (gpuCgSenseGadget.cpp:169-170 does) changes the
// operator.  As a CG regulariser it adds weight * image^2 * x.
template <class REAL>
class cuImageOperator {
public:
    cuImageOperator() { detail::require_float<REAL>(); }
    void set_weight(REAL w) { weight_ = w; }
    REAL get_weight() const { return weight_; }
    template <class A>
    void compute(A* image, bool offset_estimation = true) {
        if (!image) throw std::runtime_error("cuImageOperator::compute: illegal image");
        const auto dims = detail::dims_of(*image);
        if (!image_ || image_->get_dimensions() != dims) image_ = std::make_shared<cuNDArray<REAL>>(dims);
        float off = offset_;
        check(gt_image_operator_compute(image->get_data_ptr(), detail::elements(*image), offset_estimation ? 1 : 0,
                                        &off, image_->get_data_ptr(), nullptr));
        offset_ = off;
    }
    std::shared_ptr<cuNDArray<REAL>> get() { return image_; }
    REAL get_offset() const { return offset_; }

private:
    REAL weight_ = REAL(1);
    REAL offset_ = REAL(0);
    std::shared_ptr<cuNDArray<REAL>> image_;
};

// The CG-SENSE gadget's preconditioner weights (gpuCgSenseGadget.cpp:165-176):
// 1 / sqrt(sum_c |csm_c|^2 + kappa * reg) as complex, reg may be null.
template <class CSM, class REG>
std::shared_ptr<cuNDArray<std::complex<float>>> cg_sense_precond_weights(CSM* csm, REG* reg, float kappa) {
    if (!csm) throw std::runtime_error("cg_sense_precond_weights: illegal coil maps");
    auto dims = detail::dims_of(*csm);
    const size_t C = dims.back();
    dims.pop_back();
    auto w = std::make_shared<cuNDArray<std::complex<float>>>(dims);
    check(gt_cgsense_precond_weights(csm->get_data_ptr(), w->get_number_of_elements(), C,
                                     reg ? reinterpret_cast<const float*>(reg->get_data_ptr()) : nullptr, kappa,
                                     w->get_data_ptr(), nullptr));
    return w;
}

// cuCgSolver for a cuNonCartesianSenseOperator (cgSolver.h:105-390): solve(d)
// returns x = CG(w_E E^H E + kappa I + sum_k w_k R_k^H R_k, w_E E^H d), the
// iterations and dot products on the device.  An identity regulariser of
// weight kappa stands for add_regularization_operator(identity) with that
// weight; image regularisers are cuImageOperator objects (their current image
// and weight are read at each solve).
template <class REAL, unsigned int D>
class cuCgSolver {
public:
    void set_encoding_operator(std::shared_ptr<cuNonCartesianSenseOperator<REAL, D>> E) { E_ = std::move(E); }
    void set_max_iterations(unsigned int n) { iterations_ = n; }
    void set_tc_tolerance(REAL t) { tolerance_ = t; }
    void set_identity_regularization(REAL kappa) { kappa_ = kappa; }
    // cgSolver::add_regularization_operator (an image operator)
    void add_regularization_operator(std::shared_ptr<cuImageOperator<REAL>> R) { regs_.push_back(std::move(R)); }
    // solver::set_x0: the initial guess [N..., F] (kept by reference)
    template <class A>
    void set_x0(std::shared_ptr<A> x0) {
        x0_keep_ = x0;
        x0_ = x0 ? static_cast<const void*>(x0->get_data_ptr()) : nullptr;
    }
    // cgPreconditioner::set_weights: complex diagonal weights, kept by reference.
    template <class A>
    void set_preconditioner_weights(std::shared_ptr<A> w) {
        precond_keep_ = w;
        precond_ = w ? static_cast<const void*>(w->get_data_ptr()) : nullptr;
        precond_len_ = w ? detail::elements(*w) : 0;
    }
    // rq after each iteration of the last solve (rq_0 first).
    const std::vector<double>& get_residual_history() const { return history_; }

    template <class DATA>
    std::shared_ptr<cuNDArray<std::complex<REAL>>> solve(DATA* data) {
        if (!E_) throw std::runtime_error("cgSolver::solve: encoding operator not set");
        if (!data) throw std::runtime_error("cgSolver::solve: illegal input pointer");
        auto x = std::make_shared<cuNDArray<std::complex<REAL>>>(E_->get_domain_dimensions());
        gt_cg_desc d{};
        d.iterations = int(iterations_);
        d.tc_tolerance = tolerance_;
        d.encoding_weight = E_->get_weight();
        d.kappa = kappa_;
        d.precond = precond_;
        d.precond_len = precond_len_;
        std::vector<gt_cg_reg> regs;
        for (auto& R : regs_) {
            auto img = R->get();
            if (!img) throw std::runtime_error("cgSolver::solve: regularisation image not computed");
            regs.push_back(gt_cg_reg{double(R->get_weight()), img->get_data_ptr(), img->get_number_of_elements()});
        }
        d.regs = regs.empty() ? nullptr : regs.data();
        d.n_regs = int(regs.size());
        d.x0 = x0_;
        history_.assign(iterations_ + 1, 0.0);
        check(gt_cgsense_solve(E_->op(), data->get_data_ptr(), x->get_data_ptr(), &d, history_.data(), nullptr));
        check(gt_stream_synchronize(nullptr));
        return x;
    }

private:
    std::shared_ptr<cuNonCartesianSenseOperator<REAL, D>> E_;
    unsigned int iterations_ = 10;
    double tolerance_ = 1e-6, kappa_ = 0.0;
    std::shared_ptr<void> precond_keep_;
    const void* precond_ = nullptr;
    size_t precond_len_ = 0;
    std::vector<std::shared_ptr<cuImageOperator<REAL>>> regs_;
    std::shared_ptr<void> x0_keep_;
    const void* x0_ = nullptr;
    std::vector<double> history_;
};

}  // namespace gt_amd
