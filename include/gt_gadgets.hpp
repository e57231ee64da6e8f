// gt_gadgets.hpp -- header-only C++ restatement of the reference's radial
// CG-SENSE gadget chain and of its gridding reconstruction gadget over
// libgtnufft.so, without the Gadgetron stream framework or the ISMRMRD
// library:
//
//   gadgets/radial/gpuRadialPrepGadget.{h,cpp}     gpuRadialSensePrepGadget (modes 0-3,
//   gadgets/radial/gpuRadialSensePrepGadget.cpp     frame detection, sliding window,
//                                                   cuSenseBuffer, csm and reg images)
//   gadgets/pmri/GenericReconJob.h                 GenericReconJob
//   gadgets/pmri/gpuSenseGadget.cpp,               gpuCgSenseGadget (CG-SENSE of a job,
//   gadgets/pmri/gpuCgSenseGadget.cpp               frames out)
//   gadgets/mri_noncartesian/GriddingReconGadgetBase.hpp
//                                                  GriddingReconGadget (the DCW branch)
//
// A caller hands the gadgets what the stream would carry: an acquisition
// header (the ISMRMRD fields these gadgets read) with the profile's samples
// [coil][sample], and the job / image pairs they pass on come back as return
// values.  The configuration that process_config reads from the ISMRMRD XML
// header and the gadget properties are plain structs.  Matrix sizes are
// rounded up to multiples of warp_size as the reference does with the CUDA
// warp size; the MI355X wavefront is 64.  All GPU work goes through
// gt_pmri.hpp / gt_nufft.hpp (the library's C ABI).  Failures throw
// std::runtime_error where the reference returns GADGET_FAIL.
#pragma once

#include <algorithm>
#include <array>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <deque>
#include <memory>
#include <stdexcept>
#include <utility>
#include <vector>

#include "gt_pmri.hpp"

namespace gt_amd {

// ---------------------------------------------------------------------------
// ISMRMRD-shaped messages and configuration
// ---------------------------------------------------------------------------
namespace ismrmrd {

// ISMRMRD_ACQ_IS_NOISE_MEASUREMENT; flags are tested as bit (flag - 1).
constexpr int ACQ_IS_NOISE_MEASUREMENT = 19;
inline bool is_flag_set(uint64_t flags, int flag) { return (flags >> (flag - 1)) & 1u; }

struct EncodingCounters {
    uint16_t kspace_encode_step_1 = 0;
    uint16_t slice = 0;
    uint16_t set = 0;
};

// The acquisition header fields the gadgets read.
struct AcquisitionHeader {
    uint16_t version = 0;
    uint64_t flags = 0;
    uint32_t measurement_uid = 0;
    uint16_t number_of_samples = 0;
    uint16_t active_channels = 0;
    uint32_t acquisition_time_stamp = 0;
    uint32_t physiology_time_stamp[3] = {0, 0, 0};
    float position[3] = {0, 0, 0};
    float read_dir[3] = {0, 0, 0};
    float phase_dir[3] = {0, 0, 0};
    float slice_dir[3] = {0, 0, 0};
    float patient_table_position[3] = {0, 0, 0};
    EncodingCounters idx;
};

// The image header fields the gadgets write.
struct ImageHeader {
    uint16_t version = 0;
    uint16_t data_type = 0;
    uint32_t measurement_uid = 0;
    uint16_t matrix_size[3] = {0, 0, 0};
    float field_of_view[3] = {0, 0, 0};
    uint16_t channels = 0;
    float position[3] = {0, 0, 0};
    float read_dir[3] = {0, 0, 0};
    float phase_dir[3] = {0, 0, 0};
    float slice_dir[3] = {0, 0, 0};
    float patient_table_position[3] = {0, 0, 0};
    uint16_t slice = 0;
    uint16_t set = 0;
    uint32_t acquisition_time_stamp = 0;
    uint32_t physiology_time_stamp[3] = {0, 0, 0};
    uint16_t image_index = 0;
    uint16_t image_series_index = 0;
};
constexpr uint16_t ISMRMRD_CXFLOAT = 7;

// What process_config reads from the ISMRMRD XML header: encoding[0]'s
// encoded and recon matrix sizes, the recon field of view, the slice and set
// limits and the receiver channels.
struct EncodingHeader {
    size_t encoded_matrix[2] = {0, 0};
    size_t recon_matrix[2] = {0, 0};
    float recon_fov_mm[3] = {0, 0, 0};
    unsigned slices = 1;
    unsigned sets = 1;
    unsigned receiver_channels = 1;
};

}  // namespace ismrmrd

// A host array (hoNDArray): dimensions x fastest.
template <class T>
struct hoNDArray {
    std::vector<size_t> dims;
    std::vector<T> data;
    hoNDArray() = default;
    explicit hoNDArray(std::vector<size_t> d) : dims(std::move(d)), data(count(dims)) {}
    T* get_data_ptr() { return data.data(); }
    const T* get_data_ptr() const { return data.data(); }
    const std::vector<size_t>& get_dimensions() const { return dims; }
    size_t get_size(size_t d) const { return d < dims.size() ? dims[d] : 1; }
    size_t get_number_of_elements() const { return data.size(); }
    static size_t count(const std::vector<size_t>& d) {
        size_t n = 1;
        for (size_t v : d) n *= v;
        return n;
    }
};

namespace detail {
constexpr double kPi = 3.14159265358979323846;
template <class T>
hoNDArray<T> to_host(const cuNDArray<T>& a) {
    hoNDArray<T> h(a.get_dimensions());
    a.download(h.get_data_ptr());
    return h;
}
template <class T>
std::shared_ptr<cuNDArray<T>> to_device(const hoNDArray<T>& h) {
    auto a = std::make_shared<cuNDArray<T>>(h.dims);
    a->upload(h.get_data_ptr());
    return a;
}
inline size_t round_up(size_t v, unsigned m) { return ((v + m - 1) / m) * m; }
}  // namespace detail

// GenericReconJob (GenericReconJob.h): the arrays a SENSE job carries.
struct GenericReconJob {
    std::vector<ismrmrd::ImageHeader> image_headers;
    hoNDArray<std::complex<float>> dat;     // [samples, coils]
    hoNDArray<std::array<float, 2>> tra;    // [samples per frame, frames]
    hoNDArray<float> dcw;                   // [samples per frame]
    hoNDArray<std::complex<float>> csm;     // [N, N, coils]
    hoNDArray<std::complex<float>> reg;     // [N, N]
};

// ---------------------------------------------------------------------------
// gpuRadialSensePrepGadget
// ---------------------------------------------------------------------------

// The gadget properties (gpuRadialPrepGadget.h GADGET_PROPERTY list).
struct RadialPrepProperties {
    int mode = 0;  // 0, 1: fixed angle; 2, 3: golden ratio (original, smallest)
    int buffer_length_in_rotations = 1;
    bool buffer_using_solver = false;  // the iterative buffer: not available here
    float buffer_convolution_kernel_width = 5.5f;
    float buffer_convolution_oversampling_factor = 1.25f;
    float reconstruction_os_factor_x = 1.0f;
    float reconstruction_os_factor_y = 1.0f;
    int rotations_per_reconstruction = 0;
    int sliding_window_profiles = 0;
    int sliding_window_rotations = 0;
    int profiles_per_frame = 0;
    int frames_per_rotation = 0;
    int buffer_frames_per_rotation = 0;
    unsigned warp_size = 64;
};

class gpuRadialSensePrepGadget {
public:
    using cplx = std::complex<float>;
    using Job = std::pair<ismrmrd::ImageHeader, GenericReconJob>;

    // gpuRadialPrepGadget::process_config (gpuRadialPrepGadget.cpp:31-212)
    void process_config(const ismrmrd::EncodingHeader& h, const RadialPrepProperties& p) {
        if (p.mode < 0 || p.mode > 3) throw std::runtime_error("gpuRadialPrepGadget: illegal trajectory mode");
        if (p.buffer_using_solver)
            throw std::runtime_error("gpuRadialPrepGadget: the iterative accumulation buffer is not available");
        if (p.sliding_window_profiles > 0 && p.sliding_window_rotations > 0)
            throw std::runtime_error(
                "gpuRadialPrepGadget: sliding window over both profiles and frames is not supported");
        if (p.sliding_window_profiles > 0 && p.rotations_per_reconstruction > 0)
            throw std::runtime_error(
                "gpuRadialPrepGadget: sliding window over profiles is not supported for multiframe reconstructions");
        if (p.sliding_window_rotations > 0 && p.sliding_window_rotations >= p.rotations_per_reconstruction)
            throw std::runtime_error("gpuRadialPrepGadget: illegal sliding window configuration");
        p_ = p;
        const unsigned w = p.warp_size;
        image_dims_ = {detail::round_up(h.encoded_matrix[0], w), detail::round_up(h.encoded_matrix[1], w)};
        image_dims_recon_ = {
            detail::round_up(size_t(std::ceil(double(h.encoded_matrix[0]) * p.reconstruction_os_factor_x)), w),
            detail::round_up(size_t(std::ceil(double(h.encoded_matrix[1]) * p.reconstruction_os_factor_y)), w)};
        os_factor_ = p.buffer_convolution_oversampling_factor;
        image_dims_recon_os_ = {
            detail::round_up(size_t(std::ceil(double(image_dims_recon_[0]) * os_factor_)), w),
            detail::round_up(size_t(std::ceil(double(image_dims_recon_[1]) * os_factor_)), w)};
        os_factor_ = float(image_dims_recon_os_[0]) / float(image_dims_recon_[0]);  // the rounding kicked in
        fov_ = {h.recon_fov_mm[0], h.recon_fov_mm[1], h.recon_fov_mm[2]};
        slices_ = std::max(1u, h.slices);
        sets_ = std::max(1u, h.sets);
        streams_.clear();
        streams_.resize(size_t(slices_) * sets_);
        for (auto& s : streams_) {
            s.profiles_per_frame = p.profiles_per_frame;
            s.frames_per_rotation = p.frames_per_rotation;
            s.buffer_frames_per_rotation = p.buffer_frames_per_rotation;
            if (s.profiles_per_frame == 0) s.profiles_per_frame = long(image_dims_[0]);
            if (s.frames_per_rotation == 0)
                s.frames_per_rotation = (p.mode == 2 || p.mode == 3) ? 1 : long(image_dims_[0]) / s.profiles_per_frame;
            s.buffer = std::make_unique<cuSenseBuffer<float, 2>>();
        }
        samples_per_profile_ = -1;
        configured_ = true;
    }

    // gpuRadialPrepGadget::process (gpuRadialPrepGadget.cpp:214-544): one
    // profile, data [active_channels][number_of_samples].  Returns the job it
    // completes, if any.
    std::vector<Job> process(const ismrmrd::AcquisitionHeader& h, const cplx* data) {
        if (!configured_) throw std::runtime_error("gpuRadialPrepGadget: process before process_config");
        std::vector<Job> out;
        if (ismrmrd::is_flag_set(h.flags, ismrmrd::ACQ_IS_NOISE_MEASUREMENT)) return out;
        if (!data) throw std::runtime_error("gpuRadialPrepGadget: null profile");
        const long profile = h.idx.kspace_encode_step_1;
        const unsigned slice = h.idx.slice, set = h.idx.set;
        if (slice >= slices_ || set >= sets_) throw std::runtime_error("gpuRadialPrepGadget: slice or set out of range");
        if (samples_per_profile_ == -1) samples_per_profile_ = h.number_of_samples;
        if (samples_per_profile_ != long(h.number_of_samples))
            throw std::runtime_error("gpuRadialPrepGadget: unexpected change in the incoming profiles' lengths");
        const size_t k = size_t(set) * slices_ + slice;
        Stream& s = streams_[k];

        if (s.num_coils != h.active_channels) {  // the number of coils changed (or the first pass)
            s.num_coils = h.active_channels;
            reconfigure(k);
        }
        if (s.reconfigure) reconfigure(k);

        // the imaging plane changed: clear the accumulation buffer
        if (!same3(s.position, h.position) || !same3(s.read_dir, h.read_dir) || !same3(s.phase_dir, h.phase_dir) ||
            !same3(s.slice_dir, h.slice_dir)) {
            s.buffer->clear();
            s.buffer_update_needed = true;
            std::memcpy(s.position, h.position, sizeof s.position);
            std::memcpy(s.read_dir, h.read_dir, sizeof s.read_dir);
            std::memcpy(s.phase_dir, h.phase_dir, sizeof s.phase_dir);
            std::memcpy(s.slice_dir, h.slice_dir, sizeof s.slice_dir);
        }

        // profile numbering: profiles per frame and frames per rotation (mode dependent)
        bool new_frame_detected = false;
        if (s.previous_profile >= 0) {
            if (profile > s.previous_profile) {  // not the last profile of the frame
                if (p_.mode == 0 && p_.frames_per_rotation == 0) {
                    const long acceleration_factor = profile - s.previous_profile;
                    if (acceleration_factor != s.frames_per_rotation) {
                        s.frames_per_rotation = acceleration_factor;
                        reconfigure(k);
                    }
                }
            } else if (p_.profiles_per_frame == 0 && s.profiles_counter_frame > 0 &&
                       s.profiles_counter_frame != s.profiles_per_frame) {  // the first profile of a new frame
                new_frame_detected = true;
                s.profiles_per_frame = s.profiles_counter_frame;
                if (p_.mode == 1 && p_.frames_per_rotation == 0)
                    s.frames_per_rotation = long(image_dims_[0]) / s.profiles_per_frame;
                reconfigure(k);
            }
        }
        s.previous_profile = profile;

        const size_t S = size_t(samples_per_profile_), C = s.num_coils;
        std::vector<cplx> copy(data, data + S * C);
        if (!new_frame_detected) {
            s.frame_queue.push_back(copy);
            s.recon_queue.push_back(copy);
        }

        // the last profile of a frame updates the accumulation buffer
        bool is_last_profile_in_frame = s.profiles_counter_frame == s.profiles_per_frame - 1;
        is_last_profile_in_frame |= new_frame_detected;
        if (is_last_profile_in_frame) {
            hoNDArray<cplx> host = extract_samples(s.frame_queue, false, k);
            auto samples = detail::to_device(host);
            const long offset = s.profiles_counter_global - (new_frame_detected ? 1 : 0);
            auto traj = trajectory_for_frame(offset, k);
            s.buffer_update_needed |= s.buffer->add_frame_data(samples.get(), traj.get());
        }

        // ready to reconstruct downstream?
        long profiles_per_reconstruction = s.profiles_per_frame;
        if (p_.rotations_per_reconstruction > 0)
            profiles_per_reconstruction *= s.frames_per_rotation * p_.rotations_per_reconstruction;
        const bool is_last_profile_in_reconstruction = long(s.recon_queue.size()) == profiles_per_reconstruction;

        if (is_last_profile_in_frame || (is_last_profile_in_reconstruction && s.headers.empty())) {
            ismrmrd::ImageHeader ih;
            ih.version = h.version;
            ih.measurement_uid = h.measurement_uid;
            ih.matrix_size[0] = uint16_t(image_dims_recon_[0]);
            ih.matrix_size[1] = uint16_t(image_dims_recon_[1]);
            ih.matrix_size[2] = uint16_t(std::max(1L, s.frames_per_rotation * p_.rotations_per_reconstruction));
            for (int d = 0; d < 3; d++) ih.field_of_view[d] = fov_[size_t(d)];
            ih.channels = uint16_t(s.num_coils);
            ih.slice = h.idx.slice;
            ih.set = h.idx.set;
            ih.acquisition_time_stamp = h.acquisition_time_stamp;
            std::memcpy(ih.physiology_time_stamp, h.physiology_time_stamp, sizeof ih.physiology_time_stamp);
            std::memcpy(ih.position, h.position, sizeof ih.position);
            std::memcpy(ih.read_dir, h.read_dir, sizeof ih.read_dir);
            std::memcpy(ih.phase_dir, h.phase_dir, sizeof ih.phase_dir);
            std::memcpy(ih.slice_dir, h.slice_dir, sizeof ih.slice_dir);
            std::memcpy(ih.patient_table_position, h.patient_table_position, sizeof ih.patient_table_position);
            ih.data_type = ismrmrd::ISMRMRD_CXFLOAT;
            ih.image_index = uint16_t(s.image_counter++);
            ih.image_series_index = uint16_t(k);
            s.headers.push_back(ih);
        }

        if (is_last_profile_in_reconstruction) {
            // coil maps and the regularisation image when the buffer changed, and at the first pass
            if (s.buffer_update_needed || s.csm.data.empty() || s.reg.data.empty()) {
                auto acc = s.buffer->get_accumulated_coil_images();
                auto csm = std::make_shared<cuNDArray<cplx>>(estimate_b1_map<float, 2>(*acc));
                s.buffer->set_csm(csm);
                s.csm = detail::to_host(*csm);
                s.reg = detail::to_host(*s.buffer->get_combined_coil_image());
                s.buffer_update_needed = false;
            }
            GenericReconJob job;
            job.dat = extract_samples(s.recon_queue, true, k);
            // the trajectory follows the profiles for golden-ratio modes and frame-by-frame reconstruction
            if (p_.mode == 2 || p_.mode == 3 || p_.rotations_per_reconstruction == 0)
                trajectory_for_reconstruction(s.profiles_counter_global - (new_frame_detected ? 1 : 0), k);
            job.tra = s.traj_recon;
            job.dcw = s.dcw_recon;
            job.csm = s.csm;
            job.reg = s.reg;
            const long frames_per_reconstruction =
                std::max(1L, s.frames_per_rotation * p_.rotations_per_reconstruction);
            if (long(s.headers.size()) != frames_per_reconstruction)
                throw std::runtime_error("gpuRadialPrepGadget: unexpected size of the image header queue");
            for (long i = 0; i < frames_per_reconstruction; i++) {
                ismrmrd::ImageHeader ih = s.headers.front();
                s.headers.pop_front();
                job.image_headers.push_back(ih);
                // in sliding window mode the header goes back for reuse
                if (i >= frames_per_reconstruction - p_.sliding_window_rotations * s.frames_per_rotation)
                    s.headers.push_back(ih);
            }
            const ismrmrd::ImageHeader first = job.image_headers[0];
            out.emplace_back(first, std::move(job));
        }

        if (is_last_profile_in_frame)
            s.profiles_counter_frame = 0;
        else
            s.profiles_counter_frame++;
        if (new_frame_detected) {  // the first profile of the next frame
            s.frame_queue.push_back(copy);
            s.recon_queue.push_back(copy);
            s.profiles_counter_frame++;
        }
        s.profiles_counter_global++;
        return out;
    }

    long profiles_per_frame(unsigned set = 0, unsigned slice = 0) const {
        return streams_.at(size_t(set) * slices_ + slice).profiles_per_frame;
    }
    long frames_per_rotation(unsigned set = 0, unsigned slice = 0) const {
        return streams_.at(size_t(set) * slices_ + slice).frames_per_rotation;
    }
    std::array<size_t, 2> reconstruction_matrix() const { return image_dims_recon_; }
    std::array<size_t, 2> reconstruction_matrix_os() const { return image_dims_recon_os_; }
    float oversampling_factor() const { return os_factor_; }

private:
    struct Stream {
        long previous_profile = -1, image_counter = 0, profiles_counter_frame = 0, profiles_counter_global = 0;
        long profiles_per_frame = 0, frames_per_rotation = 0, buffer_frames_per_rotation = 0;
        unsigned num_coils = 0;
        bool buffer_update_needed = true, reconfigure = true;
        float position[3] = {0, 0, 0}, read_dir[3] = {0, 0, 0}, phase_dir[3] = {0, 0, 0}, slice_dir[3] = {0, 0, 0};
        std::deque<std::vector<cplx>> frame_queue, recon_queue;  // profiles [coil][sample]
        std::deque<ismrmrd::ImageHeader> headers;
        hoNDArray<std::array<float, 2>> traj_recon;
        hoNDArray<float> dcw_recon;
        hoNDArray<cplx> csm, reg;
        std::unique_ptr<cuSenseBuffer<float, 2>> buffer;
    };

    static bool same3(const float* a, const float* b) { return a[0] == b[0] && a[1] == b[1] && a[2] == b[2]; }
    GOLDEN_RATIO_ANGULAR_STEP_SIZE gr() const { return p_.mode == 2 ? GR_ORIGINAL : GR_SMALLEST; }
    float one_over_radial_os() const { return 1.0f / (float(samples_per_profile_) / float(image_dims_recon_[0])); }

    // extract_samples_from_queue (gpuRadialPrepGadget.cpp:798-848): [S * P, C]
    hoNDArray<cplx> extract_samples(std::deque<std::vector<cplx>>& q, bool sliding_window, size_t k) {
        const Stream& s = streams_[k];
        const size_t P = q.size(), S = size_t(samples_per_profile_), C = s.num_coils;
        hoNDArray<cplx> out(std::vector<size_t>{S * P, C});
        const long in_window = p_.sliding_window_profiles +
                               s.profiles_per_frame * s.frames_per_rotation * p_.sliding_window_rotations;
        std::deque<std::vector<cplx>> keep;
        for (size_t p = 0; p < P; p++) {
            const std::vector<cplx>& prof = q.front();
            for (size_t c = 0; c < C; c++)
                std::copy(prof.begin() + long(c * S), prof.begin() + long((c + 1) * S),
                          out.data.begin() + long(c * S * P + p * S));
            if (sliding_window && long(p) >= long(P) - in_window) keep.push_back(prof);
            q.pop_front();
        }
        for (auto& v : keep) q.push_back(std::move(v));
        return out;
    }

    // calculate_trajectory_for_reconstruction (gpuRadialPrepGadget.cpp:546-600)
    void trajectory_for_reconstruction(long offset, size_t k) {
        Stream& s = streams_[k];
        const unsigned S = unsigned(samples_per_profile_), P = unsigned(s.profiles_per_frame);
        std::shared_ptr<cuNDArray<std::array<float, 2>>> t;
        if (p_.mode == 0 || p_.mode == 1) {
            if (p_.rotations_per_reconstruction == 0) {
                const long local_frame = (offset / s.profiles_per_frame) % s.frames_per_rotation;
                const float angular_offset = float(detail::kPi) / float(s.profiles_per_frame) * float(local_frame) /
                                             float(s.frames_per_rotation);
                t = compute_radial_trajectory_fixed_angle_2d<float>(S, P, 1, angular_offset);
            } else {
                t = compute_radial_trajectory_fixed_angle_2d<float>(S, P, unsigned(s.frames_per_rotation));
            }
        } else if (p_.rotations_per_reconstruction == 0) {
            const unsigned first = unsigned(std::max(0L, offset - s.profiles_per_frame + 1));
            t = compute_radial_trajectory_golden_ratio_2d<float>(S, P, 1, first, gr());
        } else {
            const long frames = s.frames_per_rotation * p_.rotations_per_reconstruction;
            const unsigned first = unsigned(std::max(0L, offset - s.profiles_per_frame * frames + 1));
            t = compute_radial_trajectory_golden_ratio_2d<float>(S, P, unsigned(frames), first, gr());
        }
        s.traj_recon = detail::to_host(*t);
    }

    // calculate_density_compensation_for_reconstruction (gpuRadialPrepGadget.cpp:602-632)
    std::shared_ptr<cuNDArray<float>> dcw_for_frame(size_t k) const {
        const Stream& s = streams_[k];
        const unsigned S = unsigned(samples_per_profile_), P = unsigned(s.profiles_per_frame);
        if (p_.mode == 0 || p_.mode == 1)
            return compute_radial_dcw_fixed_angle_2d<float>(S, P, os_factor_, one_over_radial_os());
        return compute_radial_dcw_golden_ratio_2d<float>(S, P, os_factor_, one_over_radial_os(), 0, gr());
    }

    // calculate_trajectory_for_frame (gpuRadialPrepGadget.cpp:634-670)
    std::shared_ptr<cuNDArray<std::array<float, 2>>> trajectory_for_frame(long offset, size_t k) const {
        const Stream& s = streams_[k];
        const unsigned S = unsigned(samples_per_profile_), P = unsigned(s.profiles_per_frame);
        if (p_.mode == 0 || p_.mode == 1) {
            const long local_frame = (offset / s.profiles_per_frame) % s.frames_per_rotation;
            const float angular_offset =
                float(detail::kPi) / float(s.profiles_per_frame) * float(local_frame) / float(s.frames_per_rotation);
            return compute_radial_trajectory_fixed_angle_2d<float>(S, P, 1, angular_offset);
        }
        const unsigned first = unsigned(std::max(0L, offset - s.profiles_per_frame + 1));
        return compute_radial_trajectory_golden_ratio_2d<float>(S, P, 1, first, gr());
    }

    // gpuRadialPrepGadget::reconfigure (gpuRadialPrepGadget.cpp:861-893)
    void reconfigure(size_t k) {
        Stream& s = streams_[k];
        trajectory_for_reconstruction(0, k);
        s.dcw_recon = detail::to_host(*dcw_for_frame(k));
        s.buffer_frames_per_rotation = p_.buffer_frames_per_rotation;
        if (s.buffer_frames_per_rotation == 0)
            s.buffer_frames_per_rotation = (p_.mode == 2 || p_.mode == 3)
                                               ? long(image_dims_recon_os_[0]) / s.profiles_per_frame
                                               : s.frames_per_rotation;
        s.buffer->setup(image_dims_recon_, image_dims_recon_os_, p_.buffer_convolution_kernel_width, s.num_coils,
                        unsigned(p_.buffer_length_in_rotations), unsigned(s.buffer_frames_per_rotation));
        s.buffer->set_dcw(dcw_for_frame(k));
        s.reconfigure = false;
    }

    RadialPrepProperties p_;
    bool configured_ = false;
    std::array<size_t, 2> image_dims_{}, image_dims_recon_{}, image_dims_recon_os_{};
    float os_factor_ = 1.25f;
    std::array<float, 3> fov_{};
    unsigned slices_ = 1, sets_ = 1;
    long samples_per_profile_ = -1;
    std::vector<Stream> streams_;
};

// ---------------------------------------------------------------------------
// gpuCgSenseGadget
// ---------------------------------------------------------------------------

// The gadget properties (gpuSenseGadget.h:26-34, gpuCgSenseGadget.h:32-34).
struct CgSenseProperties {
    int setno = 0;
    int sliceno = 0;
    float oversampling_factor = 1.5f;
    float kernel_width = 5.5f;
    bool save_individual_frames = true;
    int rotations_to_discard = 0;
    float kappa = 0.3f;
    int number_of_iterations = 5;
    float cg_limit = 1e-6f;
    unsigned warp_size = 64;
};

class gpuCgSenseGadget {
public:
    using cplx = std::complex<float>;
    struct Image {
        ismrmrd::ImageHeader header;
        hoNDArray<cplx> data;  // [x, y] (one frame) or [x, y, frames]
    };

    // gpuSenseGadget::process_config, gpuCgSenseGadget::process_config
    // (gpuSenseGadget.cpp:27-62, gpuCgSenseGadget.cpp:29-82)
    void process_config(const ismrmrd::EncodingHeader& h, const CgSenseProperties& p) {
        if (p.rotations_to_discard % 2 == 1) throw std::runtime_error("gpuSenseGadget: #rotations to discard must be even");
        p_ = p;
        matrix_size_seq_ = {h.recon_matrix[0], h.recon_matrix[1]};
        channels_ = h.receiver_channels ? h.receiver_channels : 1;
        configured_ = true;
    }

    // gpuCgSenseGadget::process (gpuCgSenseGadget.cpp:84-203).  Returns false
    // (nothing done) for a job of another set or slice, which the stream
    // passes downstream.
    bool process(const ismrmrd::ImageHeader& h, const GenericReconJob& j, std::vector<Image>& out) {
        if (h.set != p_.setno || h.slice != p_.sliceno) return false;
        if (!configured_) throw std::runtime_error("gpuCgSenseGadget: data received before configuration was completed");
        if (j.csm.data.empty() || j.dat.data.empty() || j.tra.data.empty() || j.dcw.data.empty() ||
            j.reg.data.empty())
            throw std::runtime_error("gpuCgSenseGadget: received an incomplete Sense job");
        const size_t samples = j.dat.get_size(0);
        const size_t ntra = j.tra.get_number_of_elements();
        if (samples % ntra) throw std::runtime_error("gpuCgSenseGadget: samples are not a multiple of the k-space coordinates");
        const size_t rotations = samples / ntra;
        const size_t frames = j.tra.get_size(1) * rotations;
        if (rotations != 1)
            throw std::runtime_error("gpuCgSenseGadget: a job of several rotations per trajectory is not supported");

        auto traj = detail::to_device(j.tra);
        auto dcw = detail::to_device(j.dcw);
        check(gt_array_sqrt(dcw->get_data_ptr(), dcw->get_number_of_elements(), nullptr));  // sqrt for the weighting
        auto csm = detail::to_device(j.csm);
        auto data = detail::to_device(j.dat);

        matrix_size_ = {j.reg.get_size(0), j.reg.get_size(1)};
        matrix_size_os_ = {
            detail::round_up(size_t(std::ceil(double(matrix_size_[0]) * p_.oversampling_factor)), p_.warp_size),
            detail::round_up(size_t(std::ceil(double(matrix_size_[1]) * p_.oversampling_factor)), p_.warp_size)};

        auto E = std::make_shared<cuNonCartesianSenseOperator<float, 2>>();
        E->setup(matrix_size_, matrix_size_os_, p_.kernel_width);
        E->preprocess(traj.get());
        E->set_dcw(dcw);
        E->set_csm(csm);

        auto R = std::make_shared<cuImageOperator<float>>();
        R->set_weight(p_.kappa);
        auto reg = detail::to_device(j.reg);
        R->compute(reg.get());
        // preconditioner 1 / sqrt(sum |csm|^2 + kappa R): R's own image is scaled by kappa in place
        auto Rd = R->get();
        check(gt_array_scale(Rd->get_data_ptr(), Rd->get_number_of_elements(), p_.kappa, nullptr));
        auto precond = cg_sense_precond_weights(csm.get(), Rd.get(), 1.0f);

        // dcw weights on the samples
        check(gt_array_mul_weights(data->get_data_ptr(), dcw->get_data_ptr(), dcw->get_number_of_elements(),
                                   data->get_number_of_elements() / dcw->get_number_of_elements(), nullptr));
        // the samples as [M, frames, coils] for the operator
        View<cplx> d{data->get_data_ptr(), {ntra / j.tra.get_size(1), frames, j.dat.get_size(1)}};

        cuCgSolver<float, 2> cg;
        cg.set_encoding_operator(E);
        cg.add_regularization_operator(R);
        cg.set_preconditioner_weights(precond);
        cg.set_max_iterations(unsigned(p_.number_of_iterations));
        cg.set_tc_tolerance(p_.cg_limit);
        auto x = cg.solve(&d);
        hoNDArray<cplx> img = detail::to_host(*x);  // [N, N, frames]

        // crop when the recon matrix exceeds the sequence's
        std::array<size_t, 2> keep = matrix_size_;
        if (matrix_size_seq_[0] && matrix_size_seq_[1] && matrix_size_seq_ != matrix_size_) {
            keep = {std::min(matrix_size_seq_[0], matrix_size_[0]), std::min(matrix_size_seq_[1], matrix_size_[1])};
            img = crop(img, keep, frames);
        }
        put_frames(frames, rotations, j, img, keep, out);
        frame_counter_ += long(frames);
        return true;
    }

private:
    template <class T>
    struct View {
        T* p;
        std::vector<size_t> dims;
        T* get_data_ptr() const { return p; }
        const std::vector<size_t>& get_dimensions() const { return dims; }
    };

    // crop<float_complext, 2> of the centred (matrix - seq) / 2 offset: a host copy of rows
    hoNDArray<cplx> crop(const hoNDArray<cplx>& in, std::array<size_t, 2> keep, size_t frames) const {
        const size_t ox = (matrix_size_[0] - keep[0]) / 2, oy = (matrix_size_[1] - keep[1]) / 2;
        hoNDArray<cplx> out(std::vector<size_t>{keep[0], keep[1], frames});
        for (size_t f = 0; f < frames; f++)
            for (size_t y = 0; y < keep[1]; y++)
                std::copy_n(in.data.begin() + long((f * matrix_size_[1] + y + oy) * matrix_size_[0] + ox), keep[0],
                            out.data.begin() + long((f * keep[1] + y) * keep[0]));
        return out;
    }

    // gpuSenseGadget::put_frames_on_que (gpuSenseGadget.cpp:64-150)
    void put_frames(size_t frames, size_t rotations, const GenericReconJob& j, const hoNDArray<cplx>& img,
                    std::array<size_t, 2> dims, std::vector<Image>& out) const {
        size_t frames_per_rotation = frames / rotations;
        if (rotations == 1) {  // the case for golden ratio
            rotations = frames;
            frames_per_rotation = 1;
        }
        const size_t per = dims[0] * dims[1];
        if (p_.save_individual_frames) {
            for (size_t f = 0; f < frames; f++) {
                const size_t r = f / frames_per_rotation;
                const size_t half = size_t(p_.rotations_to_discard) >> 1;
                if (r < half || r >= rotations - half) continue;
                Image im;
                im.header = j.image_headers.at(f);
                im.header.matrix_size[0] = uint16_t(dims[0]);
                im.header.matrix_size[1] = uint16_t(dims[1]);
                im.header.matrix_size[2] = 1;
                im.header.channels = 1;
                im.header.image_index = uint16_t(frame_counter_ + long(f));
                im.data = hoNDArray<cplx>(std::vector<size_t>{dims[0], dims[1]});
                std::copy_n(img.data.begin() + long(f * per), per, im.data.data.begin());
                out.push_back(std::move(im));
            }
        } else {
            Image im;
            im.header = j.image_headers.at(0);
            im.header.matrix_size[0] = uint16_t(dims[0]);
            im.header.matrix_size[1] = uint16_t(dims[1]);
            im.header.matrix_size[2] = uint16_t(frames);
            im.header.channels = uint16_t(channels_);
            im.header.image_index = uint16_t(frame_counter_);
            im.data = img;
            out.push_back(std::move(im));
        }
    }

    CgSenseProperties p_;
    bool configured_ = false;
    unsigned channels_ = 1;
    std::array<size_t, 2> matrix_size_{}, matrix_size_os_{}, matrix_size_seq_{};
    long frame_counter_ = 0;
};

// ---------------------------------------------------------------------------
// GriddingReconGadget: GriddingReconGadgetBase's density-compensated branch
// ---------------------------------------------------------------------------

// process_config (GriddingReconGadgetBase.hpp:26-45) and the gridding, coil map
// and coil combination of process / reconstruct (GriddingReconGadgetBase.hpp:
// 112-132, 150-170) for one 2D encoding space.  The iterative branch (no DCW)
// is not restated.
class GriddingReconGadget {
public:
    using cplx = std::complex<float>;
    // The reference fixes the warp size at 32 here.
    void process_config(const ismrmrd::EncodingHeader& h, float kernel_width = 5.5f,
                        float gridding_oversampling_factor = 1.5f, unsigned warp_size = 32) {
        kernel_width_ = kernel_width;
        image_dims_ = {h.encoded_matrix[0], h.encoded_matrix[1]};
        image_dims_os_ = {
            detail::round_up(size_t(std::ceil(double(image_dims_[0]) * gridding_oversampling_factor)), warp_size),
            detail::round_up(size_t(std::ceil(double(image_dims_[1]) * gridding_oversampling_factor)), warp_size)};
        configured_ = true;
    }

    // data [RO, E1, CHA] (the buffer's [RO, E1, E2 = 1, CHA] with one N, S and
    // SLC), trajectory [3, RO, E1] (kx, ky, dcw) or [2, RO, E1] with dcw
    // [RO, E1] given separately.  Returns the coil-combined image [x, y].
    hoNDArray<cplx> process(const hoNDArray<cplx>& data, const hoNDArray<float>& trajectory,
                            const hoNDArray<float>* dcw_in = nullptr) {
        if (!configured_) throw std::runtime_error("GriddingReconGadget: process before process_config");
        const size_t td = trajectory.get_size(0);
        if (td != 2 && td != 3) throw std::runtime_error("GriddingReconGadget: unsupported number of trajectory dimensions");
        const size_t M = trajectory.get_number_of_elements() / td;
        if (M == 0 || data.get_number_of_elements() % M)
            throw std::runtime_error("GriddingReconGadget: data and trajectory disagree");
        const size_t C = data.get_number_of_elements() / M;
        hoNDArray<std::array<float, 2>> traj(std::vector<size_t>{M});
        hoNDArray<float> dcw(std::vector<size_t>{M});
        for (size_t i = 0; i < M; i++) {
            traj.data[i] = {trajectory.data[i * td], trajectory.data[i * td + 1]};
            if (td == 3) dcw.data[i] = trajectory.data[i * td + 2];
        }
        if (td == 2) {
            if (!dcw_in || dcw_in->get_number_of_elements() != M)
                throw std::runtime_error("GriddingReconGadget: the iterative branch (no density weights) is not available");
            dcw.data = dcw_in->data;
        }
        // dcw *= prod(image_dims_os) / asum(dcw): the scale of the input weights
        double asum = 0.0;
        for (float v : dcw.data) asum += std::fabs(double(v));
        if (!(asum > 0.0)) throw std::runtime_error("GriddingReconGadget: density weights sum to zero");
        const float scale = float(double(image_dims_os_[0] * image_dims_os_[1]) / asum);

        auto dtraj = detail::to_device(traj);
        auto ddcw = detail::to_device(dcw);
        check(gt_array_scale(ddcw->get_data_ptr(), M, scale, nullptr));
        auto ddata = detail::to_device(data);

        // gridding of every coil
        auto plan = NFFT<cuNDArray, float, 2>::make_plan(image_dims_, image_dims_os_, kernel_width_);
        plan->preprocess(*dtraj, NFFT_prep_mode::NC2C);
        View<cplx> samples{ddata->get_data_ptr(), {M, 1, C}};
        cuNDArray<cplx> images(std::vector<size_t>{image_dims_[0], image_dims_[1], C});
        plan->compute(samples, images, ddcw.get(), NFFT_comp_mode::BACKWARDS_NC2C);
        // coil maps from the coil images, then sum_c conj(csm_c) image_c (the
        // SENSE adjoint with the same weights grids the coils again and combines)
        auto csm = std::make_shared<cuNDArray<cplx>>(estimate_b1_map<float, 2>(images));
        cuNonCartesianSenseOperator<float, 2> E;
        E.setup(image_dims_, image_dims_os_, kernel_width_);
        View<std::array<float, 2>> t{dtraj->get_data_ptr(), {M, 1}};
        E.preprocess(&t);
        E.set_dcw(ddcw);
        E.set_csm(csm);
        cuNDArray<cplx> combined(std::vector<size_t>{image_dims_[0], image_dims_[1], 1});
        E.mult_MH(&samples, &combined);
        hoNDArray<cplx> out(std::vector<size_t>{image_dims_[0], image_dims_[1]});
        combined.download(out.get_data_ptr());
        return out;
    }

    std::array<size_t, 2> matrix() const { return image_dims_; }
    std::array<size_t, 2> matrix_os() const { return image_dims_os_; }

private:
    template <class T>
    struct View {
        T* p;
        std::vector<size_t> dims;
        T* get_data_ptr() const { return p; }
        const std::vector<size_t>& get_dimensions() const { return dims; }
        size_t get_number_of_elements() const { return hoNDArray<T>::count(dims); }
    };
    float kernel_width_ = 5.5f;
    std::array<size_t, 2> image_dims_{}, image_dims_os_{};
    bool configured_ = false;
};

}  // namespace gt_amd
