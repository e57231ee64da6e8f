// gadget_chain.cpp -- streams radial profiles through the gadget restatements
// of include/gt_gadgets.hpp: gpuRadialSensePrepGadget -> gpuCgSenseGadget, and
// the gridding reconstruction gadget on the first job's data.
//
// usage: gadget_chain <dir>
//   reads  <dir>/meta.txt     "N C S P F mode fpr": N x N encoded matrix, C coils,
//                             S samples per profile, P profiles per frame, F frames,
//                             trajectory mode (0 or 2) and, for mode 0, the frames
//                             per rotation of the profile numbering
//          <dir>/samples.bin  [F * P][C][S] complex64, profiles in streaming order
//   writes njobs.txt, and per job j: job<j>_{tra,dcw,dat,csm,reg,img}.bin and
//          job<j>_hdr.txt (image_index matrix_size[0..2] channels slice set of
//          the job's header, then of the output image); grid.bin: the gridding
//          gadget's coil-combined image of job 0's data (trajectory [3, S, P]
//          with job 0's dcw).  Checks that invalid use throws.  Exit status 0
//          on success.
#include <cmath>
#include <complex>
#include <fstream>
#include <functional>
#include <iostream>
#include <stdexcept>
#include <string>
#include <vector>

#include "gt_gadgets.hpp"

namespace {

using cplx = std::complex<float>;

template <class T>
void dump(const std::string& path, const std::vector<T>& v) {
    std::ofstream f(path, std::ios::binary);
    f.write(reinterpret_cast<const char*>(v.data()), std::streamsize(v.size() * sizeof(T)));
}

bool throws(const std::function<void()>& f) {
    try {
        f();
    } catch (const std::runtime_error&) {
        return true;
    }
    return false;
}

void header_line(std::ostream& o, const gt_amd::ismrmrd::ImageHeader& h) {
    o << h.image_index << " " << h.matrix_size[0] << " " << h.matrix_size[1] << " " << h.matrix_size[2] << " "
      << h.channels << " " << h.slice << " " << h.set << "\n";
}

}  // namespace

int main(int argc, char** argv) {
    if (argc != 2) {
        std::cerr << "usage: gadget_chain <dir>\n";
        return 2;
    }
    const std::string dir = argv[1];
    try {
        using namespace gt_amd;
        size_t N = 0, C = 0, S = 0, P = 0, F = 0;
        int mode = 0, fpr = 1;
        {
            std::ifstream m(dir + "/meta.txt");
            if (!(m >> N >> C >> S >> P >> F >> mode >> fpr)) throw std::runtime_error("bad meta.txt");
        }
        std::vector<cplx> samples(F * P * C * S);
        {
            std::ifstream f(dir + "/samples.bin", std::ios::binary);
            f.read(reinterpret_cast<char*>(samples.data()), std::streamsize(samples.size() * sizeof(cplx)));
            if (!f) throw std::runtime_error("bad samples.bin");
        }

        ismrmrd::EncodingHeader eh;
        eh.encoded_matrix[0] = eh.encoded_matrix[1] = N;
        eh.recon_matrix[0] = eh.recon_matrix[1] = N;
        eh.recon_fov_mm[0] = eh.recon_fov_mm[1] = 256.0f;
        eh.recon_fov_mm[2] = 8.0f;
        eh.receiver_channels = unsigned(C);

        RadialPrepProperties pp;
        pp.mode = mode;
        pp.profiles_per_frame = mode == 2 ? int(P) : 0;  // modes 0 and 1 find it from the numbering
        gpuRadialSensePrepGadget prep;
        prep.process_config(eh, pp);

        CgSenseProperties cp;
        cp.kappa = 0.3f;
        cp.number_of_iterations = 5;
        gpuCgSenseGadget cg;
        cg.process_config(eh, cp);

        std::vector<gpuRadialSensePrepGadget::Job> jobs;
        ismrmrd::AcquisitionHeader ah;
        ah.number_of_samples = uint16_t(S);
        ah.active_channels = uint16_t(C);
        ah.read_dir[0] = 1.0f;
        ah.phase_dir[1] = 1.0f;
        ah.slice_dir[2] = 1.0f;
        // a noise profile first: consumed without effect
        {
            ismrmrd::AcquisitionHeader noise = ah;
            noise.flags = uint64_t(1) << (ismrmrd::ACQ_IS_NOISE_MEASUREMENT - 1);
            if (!prep.process(noise, samples.data()).empty()) throw std::runtime_error("noise profile made a job");
        }
        for (size_t f = 0; f < F; f++)
            for (size_t p = 0; p < P; p++) {
                // mode 0: profiles k + j * fpr of rotation frame k; mode 2: 0 .. P - 1 per frame
                const size_t k = f % size_t(fpr);
                ah.idx.kspace_encode_step_1 = uint16_t(mode == 0 ? k + p * size_t(fpr) : p);
                ah.acquisition_time_stamp = uint32_t(f * P + p);
                for (auto& j : prep.process(ah, samples.data() + (f * P + p) * C * S)) jobs.push_back(std::move(j));
            }
        std::ofstream(dir + "/njobs.txt") << jobs.size() << "\n";
        std::cout << "jobs " << jobs.size() << ", profiles per frame " << prep.profiles_per_frame()
                  << ", frames per rotation " << prep.frames_per_rotation() << "\n";

        for (size_t j = 0; j < jobs.size(); j++) {
            auto& [h, job] = jobs[j];
            const std::string b = dir + "/job" + std::to_string(j);
            dump(b + "_tra.bin", job.tra.data);
            dump(b + "_dcw.bin", job.dcw.data);
            dump(b + "_dat.bin", job.dat.data);
            dump(b + "_csm.bin", job.csm.data);
            dump(b + "_reg.bin", job.reg.data);
            std::vector<gpuCgSenseGadget::Image> images;
            if (!cg.process(h, job, images)) throw std::runtime_error("the CG gadget passed its own job on");
            if (images.size() != 1) throw std::runtime_error("one image per frame expected");
            dump(b + "_img.bin", images[0].data.data);
            std::ofstream hf(b + "_hdr.txt");
            header_line(hf, h);
            header_line(hf, images[0].header);
            // a job of another set is passed on untouched
            ismrmrd::ImageHeader other = h;
            other.set = 1;
            std::vector<gpuCgSenseGadget::Image> none;
            if (cg.process(other, job, none) || !none.empty()) throw std::runtime_error("foreign job processed");
        }

        // the gridding gadget on job 0: trajectory [3, S, P] with the dcw
        {
            const auto& job = jobs.at(0).second;
            const size_t M = S * P;
            hoNDArray<float> traj(std::vector<size_t>{3, S, P});
            for (size_t i = 0; i < M; i++) {
                traj.data[3 * i] = job.tra.data[i][0];
                traj.data[3 * i + 1] = job.tra.data[i][1];
                traj.data[3 * i + 2] = job.dcw.data[i];
            }
            hoNDArray<cplx> data(std::vector<size_t>{S, P, C});
            data.data = job.dat.data;
            GriddingReconGadget grid;
            grid.process_config(eh);
            dump(dir + "/grid.bin", grid.process(data, traj).data);
            std::cout << "gridding matrix " << grid.matrix()[0] << " os " << grid.matrix_os()[0] << "\n";
        }

        // invalid use
        bool ok = true;
        gpuRadialSensePrepGadget fresh;
        ok &= throws([&] { fresh.process(ah, samples.data()); });
        RadialPrepProperties solver = pp;
        solver.buffer_using_solver = true;
        ok &= throws([&] { fresh.process_config(eh, solver); });
        ismrmrd::AcquisitionHeader shorter = ah;
        shorter.number_of_samples = uint16_t(S / 2);
        ok &= throws([&] { prep.process(shorter, samples.data()); });
        CgSenseProperties odd = cp;
        odd.rotations_to_discard = 1;
        gpuCgSenseGadget bad;
        ok &= throws([&] { bad.process_config(eh, odd); });
        if (!ok) {
            std::cerr << "an invalid call did not throw\n";
            return 1;
        }
        std::cout << "gadget chain ok\n";
    } catch (const std::exception& e) {
        std::cerr << "error: " << e.what() << "\n";
        return 1;
    }
    return 0;
}
