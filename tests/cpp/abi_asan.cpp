// Host-side checks of libgtnufft's C ABI under AddressSanitizer and
// UndefinedBehaviorSanitizer.  Built by gadgetron_amd/csrc/Makefile's `asan`
// target from the library's sources compiled for the host only (no device
// code), and run by tests/test_abi_asan.py on a host without a GPU: every call
// here takes a validation path that returns before any device work, or fails
// cleanly for want of a device.  The sanitizers turn a stray access, leak or
// undefined operation in that host code into a failing exit.
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>

#include "gt_b1map.h"
#include "gt_buffer.h"
#include "gt_cgsense.h"
#include "gt_nufft.h"
#include "gt_radial.h"

namespace {

int failures = 0;
int checks = 0;

void expect(gt_status got, gt_status want, const char* what) {
    checks++;
    if (got != want) {
        std::fprintf(stderr, "abi_asan: %s returned %d, expected %d (%s)\n", what, int(got), int(want),
                     gt_last_error());
        failures++;
    } else if (got != GT_OK && std::strlen(gt_last_error()) == 0) {
        std::fprintf(stderr, "abi_asan: %s left no message\n", what);
        failures++;
    }
}

gt_nufft_desc desc2d() {
    gt_nufft_desc d;
    std::memset(&d, 0, sizeof d);
    d.D = 2;
    d.precision = GT_FLOAT;
    d.N[0] = d.N[1] = 128;
    d.N_os[0] = d.N_os[1] = 192;
    d.W = 5.5;
    d.conv_type = GT_CONV_STANDARD;
    d.device = -1;
    return d;
}

void plan_descriptors() {
    gt_nufft_plan* p = nullptr;
    struct Case {
        const char* what;
        void (*edit)(gt_nufft_desc&);
        gt_status want;
    };
    const Case cases[] = {
        {"W below 1.8", [](gt_nufft_desc& d) { d.W = 1.5; }, GT_ERR_INVALID_ARGUMENT},
        {"W NaN", [](gt_nufft_desc& d) { d.W = 0.0 / 0.0; }, GT_ERR_INVALID_ARGUMENT},
        {"D 5", [](gt_nufft_desc& d) { d.D = 5; }, GT_ERR_UNSUPPORTED},
        {"D 0", [](gt_nufft_desc& d) { d.D = 0; }, GT_ERR_UNSUPPORTED},
        {"double ATOMIC", [](gt_nufft_desc& d) { d.precision = GT_DOUBLE; d.conv_type = GT_CONV_ATOMIC; },
         GT_ERR_UNSUPPORTED},
        {"precision 2", [](gt_nufft_desc& d) { d.precision = 2; }, GT_ERR_UNSUPPORTED},
        {"conv type 7", [](gt_nufft_desc& d) { d.conv_type = 7; }, GT_ERR_INVALID_ARGUMENT},
        {"N_os < N", [](gt_nufft_desc& d) { d.N_os[0] = 100; }, GT_ERR_INVALID_ARGUMENT},
        {"odd N_os", [](gt_nufft_desc& d) { d.N_os[0] = 193; }, GT_ERR_UNSUPPORTED},
        {"os factor 0.5", [](gt_nufft_desc& d) { d.os_factor = 0.5; }, GT_ERR_INVALID_ARGUMENT},
        {"W 9.5", [](gt_nufft_desc& d) { d.W = 9.5; }, GT_ERR_UNSUPPORTED},
        {"N 0", [](gt_nufft_desc& d) { d.N[1] = 0; }, GT_ERR_INVALID_ARGUMENT},
    };
    for (const Case& c : cases) {
        gt_nufft_desc d = desc2d();
        c.edit(d);
        expect(gt_nufft_create(&p, &d), c.want, c.what);
        if (p) {
            std::fprintf(stderr, "abi_asan: %s returned a plan\n", c.what);
            failures++;
        }
    }
    // a valid descriptor: no device here, so it fails cleanly (or succeeds on a GPU host)
    gt_nufft_desc d = desc2d();
    const gt_status st = gt_nufft_create(&p, &d);
    checks++;
    if (st == GT_OK) {
        expect(gt_nufft_destroy(p), GT_OK, "destroy");
    } else if (st != GT_ERR_DEVICE || p) {
        std::fprintf(stderr, "abi_asan: valid create returned %d\n", int(st));
        failures++;
    }
}

void null_arguments() {
    const gt_nufft_desc d = desc2d();
    gt_nufft_plan* p = nullptr;
    size_t dims[3] = {1, 1, 1};
    expect(gt_nufft_create(nullptr, &d), GT_ERR_INVALID_ARGUMENT, "create null out");
    expect(gt_nufft_create(&p, nullptr), GT_ERR_INVALID_ARGUMENT, "create null desc");
    expect(gt_nufft_destroy(nullptr), GT_OK, "destroy null");
    expect(gt_nufft_preprocess(nullptr, nullptr, 1, 1, GT_PREP_ALL, nullptr), GT_ERR_INVALID_ARGUMENT,
           "preprocess");
    expect(gt_nufft_compute(nullptr, nullptr, dims, 3, nullptr, dims, 3, nullptr, 0, 0, nullptr),
           GT_ERR_INVALID_ARGUMENT, "compute");
    expect(gt_nufft_compute_noalloc(nullptr, nullptr, dims, 3, nullptr, dims, 3, nullptr, 0, 0, nullptr),
           GT_ERR_INVALID_ARGUMENT, "compute_noalloc");
    expect(gt_nufft_mult_MH_M(nullptr, nullptr, nullptr, dims, 3, nullptr, 0, nullptr), GT_ERR_INVALID_ARGUMENT,
           "mult_MH_M");
    expect(gt_nufft_convolve(nullptr, nullptr, dims, 3, nullptr, dims, 3, 0, 0, nullptr), GT_ERR_INVALID_ARGUMENT,
           "convolve");
    expect(gt_nufft_fft(nullptr, nullptr, dims, 3, 0, 1, nullptr), GT_ERR_INVALID_ARGUMENT, "fft");
    expect(gt_nufft_deapodize(nullptr, nullptr, dims, 3, 0, nullptr), GT_ERR_INVALID_ARGUMENT, "deapodize");
    size_t N[4];
    double W = 0.0;
    expect(gt_nufft_get_matrix_size(nullptr, N), GT_ERR_INVALID_ARGUMENT, "get_matrix_size");
    expect(gt_nufft_get_W(nullptr, &W), GT_ERR_INVALID_ARGUMENT, "get_W");
    expect(gt_nufft_profile(nullptr, 1), GT_ERR_INVALID_ARGUMENT, "profile");
    expect(gt_nufft_profile_stages(nullptr, 3u), GT_ERR_INVALID_ARGUMENT, "profile_stages");
    expect(gt_array_add(nullptr, nullptr, 0, GT_FLOAT, nullptr), GT_OK, "empty array_add");
    expect(gt_array_add(nullptr, nullptr, 4, 7, nullptr), GT_ERR_INVALID_ARGUMENT, "array_add precision");
    expect(gt_array_add(nullptr, nullptr, 4, GT_FLOAT, nullptr), GT_ERR_INVALID_ARGUMENT, "array_add null");

    gt_sense_op* op = nullptr;
    expect(gt_sense_create(&op, nullptr, nullptr, 4, nullptr, 0), GT_ERR_INVALID_ARGUMENT, "sense_create");
    expect(gt_sense_destroy(nullptr), GT_OK, "sense_destroy null");
    expect(gt_sense_mult_M(nullptr, nullptr, nullptr, 0, nullptr), GT_ERR_INVALID_ARGUMENT, "sense_mult_M");
    gt_cg_desc cg;
    std::memset(&cg, 0, sizeof cg);
    cg.iterations = 10;
    expect(gt_cgsense_solve(nullptr, nullptr, nullptr, &cg, nullptr, nullptr), GT_ERR_INVALID_ARGUMENT,
           "cgsense_solve");
    gt_cg_batch* batch = nullptr;
    expect(gt_cgsense_batch_create(&batch, nullptr, nullptr, nullptr, 0, &cg, GT_CG_BATCHED, nullptr),
           GT_ERR_INVALID_ARGUMENT, "batch_create");
    expect(gt_cgsense_batch_run(nullptr, nullptr), GT_ERR_INVALID_ARGUMENT, "batch_run");
    float off = 0.0f;
    expect(gt_image_operator_compute(nullptr, 16, 1, &off, nullptr, nullptr), GT_ERR_INVALID_ARGUMENT,
           "image_operator_compute");
    expect(gt_cgsense_precond_weights(nullptr, 16, 4, nullptr, 0.0, nullptr, nullptr), GT_ERR_INVALID_ARGUMENT,
           "precond_weights");

    expect(gt_radial_trajectory_golden_ratio_2d(nullptr, 64, 8, 1, 0, 0, nullptr), GT_ERR_INVALID_ARGUMENT,
           "golden ratio null");
    expect(gt_radial_trajectory_fixed_angle_2d(nullptr, 64, 8, 1, 0.0f, nullptr), GT_ERR_INVALID_ARGUMENT,
           "fixed angle null");
    expect(gt_kooshball_trajectory_3d(nullptr, 64, 8, nullptr, nullptr), GT_ERR_INVALID_ARGUMENT, "kooshball null");
    size_t ms[2] = {64, 64};
    expect(gt_sdc_estimate_dcw(nullptr, 16, 2, ms, 2.1f, 10, 5.5f, nullptr, nullptr, nullptr),
           GT_ERR_INVALID_ARGUMENT, "sdc null");

    gt_buffer* buf = nullptr;
    expect(gt_buffer_create(nullptr), GT_ERR_INVALID_ARGUMENT, "buffer_create null");
    expect(gt_buffer_create(&buf), GT_OK, "buffer_create");
    expect(gt_buffer_setup(buf, 5, ms, ms, 5.5f, 4, 1, 1), GT_ERR_UNSUPPORTED, "buffer_setup D 5");
    float f = 0.0f;
    expect(gt_buffer_get_normalization_factor(nullptr, &f), GT_ERR_INVALID_ARGUMENT, "normalization null");
    expect(gt_buffer_destroy(buf), GT_OK, "buffer_destroy");
    expect(gt_estimate_b1_map(nullptr, 2, ms, 4, -1, nullptr, nullptr), GT_ERR_INVALID_ARGUMENT, "b1 null");
}

// gt_last_error is per thread: two threads' failures keep their own messages.
void thread_local_messages() {
    std::string a, b;
    std::thread t1([&] {
        gt_nufft_desc d = desc2d();
        d.W = 1.0;
        gt_nufft_plan* p = nullptr;
        (void)gt_nufft_create(&p, &d);
        a = gt_last_error();
    });
    std::thread t2([&] {
        gt_nufft_desc d = desc2d();
        d.D = 9;
        gt_nufft_plan* p = nullptr;
        (void)gt_nufft_create(&p, &d);
        b = gt_last_error();
    });
    t1.join();
    t2.join();
    checks++;
    if (a.empty() || b.empty() || a == b) {
        std::fprintf(stderr, "abi_asan: thread messages '%s' / '%s'\n", a.c_str(), b.c_str());
        failures++;
    }
}

}  // namespace

int main() {
    plan_descriptors();
    null_arguments();
    thread_local_messages();
    if (failures) {
        std::fprintf(stderr, "abi_asan: %d of %d checks failed\n", failures, checks);
        return 1;
    }
    std::printf("abi_asan: %d checks ok\n", checks);
    return 0;
}
