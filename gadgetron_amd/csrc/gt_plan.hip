// gt_plan.hip -- host side of libgtnufft: the NFFT plan behind the C ABI of
// include/gt_nufft.h.
//
// Mirrors NFFT_plan<cuNDArray,float,D> (toolboxes/nfft/NFFT.hpp) with the
// cuNFFT_impl specifics (toolboxes/nfft/gpu/cuNFFT.cu) and the gridding
// convolution object (toolboxes/nfft/gpu/cuGriddingConvolution.cu), re-planned
// for MI355X:
//  * preprocess bins the samples into grid tiles on the device (key emission,
//    scan, stable radix sort), splits dense tiles into work items and, for the
//    2D hot path, writes the per-sample Kaiser-Bessel tables the gridding
//    kernels read with uniform loads -- once per trajectory;
//  * a 2D compute keeps its oversampled grids in the row-interleaved layout
//    [y][batch][x], so each FFT is two single rocFFT passes (rows, then
//    columns with batch distance 1) instead of a plan per image, the rows that
//    are known zero (forward) or cropped away (adjoint) are skipped, and pad /
//    crop / deapodization / centring / scaling run as one row kernel each;
//  * rocFFT plans are created once per (pass, batch) and cached;
//  * workspaces grow on demand and are kept across calls.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>
#include <rocfft/rocfft.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <sstream>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../../include/gt_nufft.h"
#include "gt_common.h"
#include "gt_kb.h"
#include "gt_kernels.h"
#include "gt_kernels_fast.h"
#include "gt_kernels_reg.h"
#include "gt_kernels_3d.h"
#include "gt_kernels_d.h"
#include "gt_fft.h"

std::string& gt_host::last_error() {
    thread_local std::string message;
    return message;
}

namespace {

using namespace gt_host;

std::once_flag g_rocfft_once;

struct FFTPlan {
    rocfft_plan fwd = nullptr, bwd = nullptr;
    rocfft_plan fwd2 = nullptr, bwd2 = nullptr;  // D = 4: the 1-D pass along the 4th dimension
    rocfft_execution_info info = nullptr;
    DevBuf work;
};

#define GT_FFT(call)                                                           \
    do {                                                                       \
        rocfft_status s_ = (call);                                             \
        if (s_ != rocfft_status_success) {                                     \
            std::ostringstream os_;                                            \
            os_ << #call << " failed with rocfft status " << int(s_);          \
            throw gt_error(GT_ERR_FFT, os_.str());                             \
        }                                                                      \
    } while (0)

enum FFTKind { FFT_IMAGE = 0, FFT_ROWS = 1, FFT_COLS = 2, FFT_GRID2D = 3 };

// How a 2D compute transforms its row-interleaved grids:
//   2 (default) the fused band transforms of gt_fft.h (pad / crop and
//     deapodization inside the row transforms), for compiled lengths;
//   1 one strided 2-D rocFFT plan; 0 two 1-D rocFFT passes that skip the zero
//     or cropped rows (1 and 0 with separate pad / crop kernels).
// The GT_NUFFT_FFT environment variable ("fused", "grid2d", "passes")
// overrides the default.
int fft2d_strategy() {
    static int mode = [] {
        const char* e = std::getenv("GT_NUFFT_FFT");
        if (e && std::string(e) == "passes") return 0;
        if (e && std::string(e) == "grid2d") return 1;
        return 2;
    }();
    return mode;
}

// Transform lengths with compiled fused kernels.
template <template <int> class K, class... A>
bool dispatch_fft_len(int L, A&&... a) {
    switch (L) {
#define GT_LEN(l) \
    case l: K<l>::run(std::forward<A>(a)...); return true;
        GT_LEN(96) GT_LEN(128) GT_LEN(144) GT_LEN(160) GT_LEN(192) GT_LEN(240) GT_LEN(256) GT_LEN(288)
        GT_LEN(320) GT_LEN(384)
#undef GT_LEN
    default: return false;
    }
}
bool fused_fft_len(int L) {
    switch (L) {
    case 96: case 128: case 144: case 160: case 192: case 240: case 256: case 288: case 320: case 384:
        return true;
    default: return false;
    }
}

// Sample lists of one binning: the tile-sorted sample indices, the
// matching keys, and the work items over them.
struct Binning {
    gt::GridGeom geom{};
    DevBuf list;      // sample index per entry, sorted by tile key then sample
    DevBuf table;     // per-entry Kaiser-Bessel table (2D hot path)
    DevBuf heads;     // per-entry (sample index, row mask) of the NC2C table
    DevBuf dcw_e;     // dcw per entry (3D C2NC, per-sample dcw)
    DevBuf wtab;      // per-entry weight rows of the matrix-core 2D C2NC
    DevBuf items;     // gt::WorkItem
    DevBuf red;       // gt::RedItem (split tiles)
    DevBuf recs;      // work-item records of the matrix-core 3D C2NC
    size_t n = 0;
    int n_items = 0, n_red = 0, n_partials = 0;
    void release() {
        list.release();
        dcw_e.release();
        wtab.release();
        table.release();
        heads.release();
        items.release();
        red.release();
        recs.release();
        n = 0;
        n_items = n_red = n_partials = 0;
    }
};

}  // namespace

struct gt_nufft_plan {
    int D = 2;
    size_t N[4] = {1, 1, 1, 1}, Nos[4] = {1, 1, 1, 1};
    double W = 0;
    float Wf = 0, radius = 0, one_over_W = 0, beta[4] = {0, 0, 0, 0};
    int conv_type = GT_CONV_STANDARD;
    int device = 0;
    int cus = 0;  // compute units of the device
    int KW = 0;
    int tile_req[4] = {0, 0, 0, 0};
    int max_items_req = 0;
    int nc2c_len_req = 0;  // NC2C entries per work item chosen by a fused batch's tuner (0: the rule below)
    bool fast2d = false;  // the 2D hot path (tables, row-interleaved grids)
    bool fft_fused = false;  // D == 2 with both grid lengths among the fused FFT lengths
    bool fft_fused3 = false; // D == 3 (float) with every grid length among the fused FFT lengths
    bool fast3d = false;     // the 3D hot path (register-tile NC2C, staged C2NC)
    bool c2nc3_mfma = false; // fast3d with the matrix-core C2NC
    int c2nc3_ey = 8;        // its staged rows in y (8, 12 or 16)
    bool dbl = false;        // GT_DOUBLE: the generic path in double precision
    double radius_d = 0, one_over_W_d = 0, beta_d[4] = {0, 0, 0, 0}, nos_d[4] = {1, 1, 1, 1};

    gt::GridGeom geom{};  // geometry for the generic kernels
    std::vector<float> filt_host[4];
    DevBuf filt[4];
    DevBuf filt_d[4];   // GT_DOUBLE: the filters in double
    DevBuf twiddle[3];  // exp(-2 pi i m / N_os[d]), m < N_os[d] (fused FFT)

    // preprocess state
    size_t M = 0, F = 0;
    bool prep_c2nc = false, prep_nc2c = false;
    // generation: bumped by preprocess, set_tiling and every reallocation of
    // a buffer below (their epoch points here), so a recorded launch sequence
    // can detect that it refers to stale state
    uint64_t generation = 0;
    DevBuf traj;  // scaled positions [M*F][D] (GT_DOUBLE: a float copy for the binning)
    DevBuf traj_d;  // GT_DOUBLE: scaled positions in double
    DevBuf row_inv;     // SPARSE_MATRIX: 1 / window weight sum per sample [M*F]
    DevBuf sparse_dcw;  // SPARSE_MATRIX: caller's dcw times row_inv
    DevBuf flag;
    Binning bn;   // NC2C: (tile, sample) pairs for every tile a window touches
    Binning bc;   // C2NC: one entry per sample in its home tile (2D hot path)
    Binning bm;   // 2D matrix-core C2NC: home tiles whose halo fills 16-cell staged rows

    // workspaces
    DevBuf grid;
    DevBuf tmp_samples;
    DevBuf partials;

    std::map<std::pair<int, size_t>, FFTPlan> ffts;

    // stage timing (gt_nufft_profile, gt_nufft_profile_stages)
    bool profiling = false;
    unsigned stage_mask = 0;  // stages that record events while profiling
    struct Record {
        int stage;
        hipEvent_t a, b;
    };
    std::vector<Record> records;
    std::vector<hipEvent_t> event_pool;
    hipEvent_t get_event() {
        if (!event_pool.empty()) {
            hipEvent_t e = event_pool.back();
            event_pool.pop_back();
            return e;
        }
        hipEvent_t e = nullptr;
        if (hipEventCreate(&e) != hipSuccess) return nullptr;
        return e;
    }

    ~gt_nufft_plan() {
        for (auto& r : records) {
            (void)hipEventDestroy(r.a);
            (void)hipEventDestroy(r.b);
        }
        for (auto e : event_pool) (void)hipEventDestroy(e);
        for (auto& kv : ffts) {
            if (kv.second.fwd) rocfft_plan_destroy(kv.second.fwd);
            if (kv.second.bwd) rocfft_plan_destroy(kv.second.bwd);
            if (kv.second.fwd2) rocfft_plan_destroy(kv.second.fwd2);
            if (kv.second.bwd2) rocfft_plan_destroy(kv.second.bwd2);
            if (kv.second.info) rocfft_execution_info_destroy(kv.second.info);
            kv.second.work.release();
        }
        for (auto& b : filt) b.release();
        for (auto& b : filt_d) b.release();
        traj_d.release();
        traj.release();
        flag.release();
        bn.release();
        bc.release();
        bm.release();
        grid.release();
        tmp_samples.release();
        partials.release();
    }

    // points every workspace and table buffer's epoch at `generation`
    void track() {
        for (DevBuf* b : {&traj, &traj_d, &row_inv, &sparse_dcw, &flag, &grid, &tmp_samples, &partials})
            b->epoch = &generation;
        for (Binning* bb : {&bn, &bc, &bm})
            for (DevBuf* b : {&bb->list, &bb->table, &bb->heads, &bb->dcw_e, &bb->wtab, &bb->items, &bb->red,
                              &bb->recs})
                b->epoch = &generation;
    }

    size_t G() const {
        size_t g = 1;
        for (int d = 0; d < D; d++) g *= Nos[d];
        return g;
    }
    size_t Nimg() const {
        size_t n = 1;
        for (int d = 0; d < D; d++) n *= N[d];
        return n;
    }
};

namespace {

// Events gt_nufft_profile keeps ready (8 per adjoint+forward pair of the 2D
// path: 128 pairs between reads before the pool has to grow).
constexpr size_t kEventPoolSize = 1024;

// Records a start/stop event pair around one stage when the plan profiles.
struct StageTimer {
    gt_nufft_plan* p;
    int stage;
    hipStream_t s;
    hipEvent_t a = nullptr;
    StageTimer(gt_nufft_plan* p_, int stage_, hipStream_t s_) : p(p_), stage(stage_), s(s_) {
        if (p->profiling && ((p->stage_mask >> stage) & 1u)) {
            a = p->get_event();
            if (a && hipEventRecord(a, s) != hipSuccess) {
                p->event_pool.push_back(a);
                a = nullptr;
            }
        }
    }
    ~StageTimer() {
        if (!a) return;
        hipEvent_t b = p->get_event();
        if (b && hipEventRecord(b, s) == hipSuccess) {
            p->records.push_back({stage, a, b});
        } else {
            p->event_pool.push_back(a);
            if (b) p->event_pool.push_back(b);
        }
    }
};

// 1-D reciprocal deapodization filter along one dimension: the centred,
// unscaled DFT of the kernel sampled at x - N_os/2 (cuNFFT.cu:135-202,
// hoNFFT.cpp:84-197).  The kernel samples are symmetric about N_os/2 and
// N_os is even, so the transform is real and direction independent; the
// D-dimensional filter is the product of these 1-D filters.
std::vector<float> deapod_1d(size_t nos, float radius, float one_over_W, float beta) {
    std::vector<double> h(nos);
    for (size_t x = 0; x < nos; x++) {
        float off = float(int(x)) - float(nos) / 2.0f;
        float r = std::fabs(off);
        h[x] = (r > radius) ? 0.0 : double(gt::kb1(r, float(nos), one_over_W, beta));
    }
    std::vector<float> out(nos);
    const double c = double(nos / 2);
    for (size_t k = 0; k < nos; k++) {
        double acc = 0.0;
        for (size_t x = 0; x < nos; x++) {
            if (h[x] == 0.0) continue;
            acc += h[x] * std::cos(2.0 * M_PI * (double(k) - c) * (double(x) - c) / double(nos));
        }
        out[k] = float(1.0 / acc);
    }
    return out;
}

// The same filter in double (GT_DOUBLE plans).
std::vector<double> deapod_1d_d(size_t nos, double radius, double one_over_W, double beta) {
    std::vector<double> h(nos);
    for (size_t x = 0; x < nos; x++) {
        const double r = std::fabs(double(int(x)) - double(nos) / 2.0);
        h[x] = (r > radius) ? 0.0 : gt::kb1d(r, double(nos), one_over_W, beta);
    }
    std::vector<double> out(nos);
    const double c = double(nos / 2);
    for (size_t k = 0; k < nos; k++) {
        double acc = 0.0;
        for (size_t x = 0; x < nos; x++) {
            if (h[x] == 0.0) continue;
            acc += h[x] * std::cos(2.0 * M_PI * (double(k) - c) * (double(x) - c) / double(nos));
        }
        out[k] = 1.0 / acc;
    }
    return out;
}

void set_device(const gt_nufft_plan* p) { GT_HIP(hipSetDevice(p->device)); }

void finish_fft_plan(FFTPlan& fp) {
    size_t w1 = 0, w2 = 0, w3 = 0, w4 = 0;
    GT_FFT(rocfft_plan_get_work_buffer_size(fp.fwd, &w1));
    GT_FFT(rocfft_plan_get_work_buffer_size(fp.bwd, &w2));
    if (fp.fwd2) GT_FFT(rocfft_plan_get_work_buffer_size(fp.fwd2, &w3));
    if (fp.bwd2) GT_FFT(rocfft_plan_get_work_buffer_size(fp.bwd2, &w4));
    GT_FFT(rocfft_execution_info_create(&fp.info));
    const size_t w = std::max(std::max(w1, w2), std::max(w3, w4));
    if (w) {
        fp.work.reserve(w);
        GT_FFT(rocfft_execution_info_set_work_buffer(fp.info, fp.work.ptr, w));
    }
}

// FFT_IMAGE: D-dim transform of each image of the caller's layout [x, y, z, batch];
// FFT_ROWS:  1-D transforms of contiguous rows of length N_os[0];
// FFT_COLS:  1-D transforms along y of the [y][batch][x] layout (stride batch*N_os[0],
//            batch distance 1), `batch` = number of images.
FFTPlan& get_fft(gt_nufft_plan* p, int kind, size_t batch) {
    const auto key = std::make_pair(kind, batch);
    auto it = p->ffts.find(key);
    if (it != p->ffts.end()) return it->second;
    std::call_once(g_rocfft_once, [] { rocfft_setup(); });
    FFTPlan fp;
    if (kind == FFT_IMAGE && p->D == 4) {
        // rocFFT transforms have rank <= 3: the 3-D transforms of every volume
        // (batch * N_os[3] of them), then 1-D transforms along the 4th dimension
        // (stride V = N_os[0] N_os[1] N_os[2], V of them per image)
        const auto prec = p->dbl ? rocfft_precision_double : rocfft_precision_single;
        size_t lengths[3] = {p->Nos[0], p->Nos[1], p->Nos[2]};
        GT_FFT(rocfft_plan_create(&fp.fwd, rocfft_placement_inplace, rocfft_transform_type_complex_forward, prec, 3,
                                  lengths, batch * p->Nos[3], nullptr));
        GT_FFT(rocfft_plan_create(&fp.bwd, rocfft_placement_inplace, rocfft_transform_type_complex_inverse, prec, 3,
                                  lengths, batch * p->Nos[3], nullptr));
        rocfft_plan_description desc = nullptr;
        GT_FFT(rocfft_plan_description_create(&desc));
        size_t len = p->Nos[3], stride = p->Nos[0] * p->Nos[1] * p->Nos[2];
        const size_t dist = 1;
        rocfft_status st = rocfft_plan_description_set_data_layout(
            desc, rocfft_array_type_complex_interleaved, rocfft_array_type_complex_interleaved, nullptr, nullptr, 1,
            &stride, dist, 1, &stride, dist);
        if (st == rocfft_status_success)
            st = rocfft_plan_create(&fp.fwd2, rocfft_placement_inplace, rocfft_transform_type_complex_forward, prec, 1,
                                    &len, stride, desc);
        if (st == rocfft_status_success)
            st = rocfft_plan_create(&fp.bwd2, rocfft_placement_inplace, rocfft_transform_type_complex_inverse, prec, 1,
                                    &len, stride, desc);
        rocfft_plan_description_destroy(desc);
        GT_FFT(st);
    } else if (kind == FFT_IMAGE) {
        size_t lengths[3] = {p->Nos[0], p->Nos[1], p->Nos[2]};  // fastest dimension first
        GT_FFT(rocfft_plan_create(&fp.fwd, rocfft_placement_inplace, rocfft_transform_type_complex_forward,
                                  (p->dbl ? rocfft_precision_double : rocfft_precision_single), p->D, lengths, batch, nullptr));
        GT_FFT(rocfft_plan_create(&fp.bwd, rocfft_placement_inplace, rocfft_transform_type_complex_inverse,
                                  (p->dbl ? rocfft_precision_double : rocfft_precision_single), p->D, lengths, batch, nullptr));
    } else if (kind == FFT_GRID2D) {
        // 2-D transforms of the [y][batch][x] layout: x stride 1, y stride batch*N_os[0],
        // one transform per batch at distance N_os[0]
        rocfft_plan_description desc = nullptr;
        GT_FFT(rocfft_plan_description_create(&desc));
        size_t lengths[2] = {p->Nos[0], p->Nos[1]};
        size_t strides[2] = {1, batch * p->Nos[0]};
        const size_t dist = p->Nos[0];
        rocfft_status st = rocfft_plan_description_set_data_layout(
            desc, rocfft_array_type_complex_interleaved, rocfft_array_type_complex_interleaved, nullptr,
            nullptr, 2, strides, dist, 2, strides, dist);
        if (st == rocfft_status_success)
            st = rocfft_plan_create(&fp.fwd, rocfft_placement_inplace, rocfft_transform_type_complex_forward,
                                    (p->dbl ? rocfft_precision_double : rocfft_precision_single), 2, lengths, batch, desc);
        if (st == rocfft_status_success)
            st = rocfft_plan_create(&fp.bwd, rocfft_placement_inplace, rocfft_transform_type_complex_inverse,
                                    (p->dbl ? rocfft_precision_double : rocfft_precision_single), 2, lengths, batch, desc);
        rocfft_plan_description_destroy(desc);
        GT_FFT(st);
    } else {
        rocfft_plan_description desc = nullptr;
        GT_FFT(rocfft_plan_description_create(&desc));
        size_t len, stride, dist, count;
        if (kind == FFT_ROWS) {
            len = p->Nos[0];
            stride = 1;
            dist = p->Nos[0];
            count = batch;
        } else {
            len = p->Nos[1];
            stride = batch * p->Nos[0];
            dist = 1;
            count = batch * p->Nos[0];
        }
        rocfft_status st = rocfft_plan_description_set_data_layout(
            desc, rocfft_array_type_complex_interleaved, rocfft_array_type_complex_interleaved, nullptr,
            nullptr, 1, &stride, dist, 1, &stride, dist);
        if (st == rocfft_status_success)
            st = rocfft_plan_create(&fp.fwd, rocfft_placement_inplace, rocfft_transform_type_complex_forward,
                                    (p->dbl ? rocfft_precision_double : rocfft_precision_single), 1, &len, count, desc);
        if (st == rocfft_status_success)
            st = rocfft_plan_create(&fp.bwd, rocfft_placement_inplace, rocfft_transform_type_complex_inverse,
                                    (p->dbl ? rocfft_precision_double : rocfft_precision_single), 1, &len, count, desc);
        rocfft_plan_description_destroy(desc);
        GT_FFT(st);
    }
    finish_fft_plan(fp);
    return p->ffts.emplace(key, std::move(fp)).first->second;
}

void exec_fft(gt_nufft_plan* p, int kind, size_t batch, void* data, bool forward, hipStream_t s) {
    FFTPlan& fp = get_fft(p, kind, batch);
    GT_FFT(rocfft_execution_info_set_stream(fp.info, s));
    void* buf[1] = {data};
    GT_FFT(rocfft_execute(forward ? fp.fwd : fp.bwd, buf, nullptr, fp.info));
    if (fp.fwd2) {
        const size_t esz = p->dbl ? sizeof(double2) : sizeof(float2);
        for (size_t b = 0; b < batch; b++) {
            void* img[1] = {static_cast<char*>(data) + b * p->G() * esz};
            GT_FFT(rocfft_execute(forward ? fp.fwd2 : fp.bwd2, img, nullptr, fp.info));
        }
    }
}

void run_fft(gt_nufft_plan* p, void* data, size_t batch, bool forward, hipStream_t s) {
    StageTimer timer(p, GT_STAGE_FFT, s);
    exec_fft(p, FFT_IMAGE, batch, data, forward, s);
}

unsigned grid_blocks(long long total, int threads = 256) {
    long long b = (total + threads - 1) / threads;
    return (unsigned)std::max<long long>(1, std::min<long long>(b, 65536LL * 8));
}

float centring_global_sign(const gt_nufft_plan* p) {
    int s = 0;
    for (int d = 0; d < p->D; d++) s += int(p->Nos[d] / 2);
    return (s & 1) ? -1.0f : 1.0f;
}

gt::PadCrop padcrop(const gt_nufft_plan* p, bool pad_side, bool in_os, float scale, int sign) {
    gt::PadCrop pc{};
    for (int d = 0; d < 4; d++) {
        pc.N[d] = d < p->D ? int(p->N[d]) : 1;
        if (d < p->D) {
            // pad offset N_os/2 - N/2 (cuNDArray_utils.cu:283-286), crop offset (N_os-N)>>1 (NFFT.hpp:178)
            pc.off[d] = pad_side ? int(p->Nos[d] / 2 - p->N[d] / 2) : int((p->Nos[d] - p->N[d]) >> 1);
            pc.filt[d] = p->filt[d].as<float>();
        } else {
            pc.off[d] = 0;
            pc.filt[d] = nullptr;
        }
    }
    pc.scale = scale;
    pc.sign = sign;
    pc.in_os = in_os ? 1 : 0;
    return pc;
}

gt::Layout caller_layout(const gt_nufft_plan* p) {
    gt::Layout L;
    L.sy = (long long)p->Nos[0];
    L.sz = (long long)(p->Nos[0] * p->Nos[1]);
    L.sb = (long long)p->G();
    return L;
}

gt::Layout internal_layout(const gt_nufft_plan* p, size_t nb) {
    gt::Layout L;
    L.sb = (long long)p->Nos[0];
    L.sy = (long long)(nb * p->Nos[0]);
    L.sz = (long long)(nb * p->Nos[0] * p->Nos[1]);
    return L;
}

// ---- kernel dispatch over (D, KW) ----------------------------------------
template <template <int, int> class L, class... A>
void dispatch(int D, int KW, A&&... a) {
#define GT_CASE(d, k) \
    if (D == d && KW == k) { L<d, k>::run(std::forward<A>(a)...); return; }
#define GT_CASES(d) GT_CASE(d, 2) GT_CASE(d, 3) GT_CASE(d, 4) GT_CASE(d, 5) GT_CASE(d, 6) \
    GT_CASE(d, 7) GT_CASE(d, 8) GT_CASE(d, 9)
    GT_CASES(1)
    GT_CASES(2)
    GT_CASES(3)
#undef GT_CASES
#undef GT_CASE
    throw gt_error(GT_ERR_UNSUPPORTED, "unsupported (D, window) combination");
}

// dispatch for the kernels that also take D = 4 (the generic gridding kernels)
template <template <int, int> class L, class... A>
void dispatch4(int D, int KW, A&&... a) {
    if (D <= 3) {
        dispatch<L>(D, KW, std::forward<A>(a)...);
        return;
    }
#define GT_CASE(k) \
    if (D == 4 && KW == k) { L<4, k>::run(std::forward<A>(a)...); return; }
    GT_CASE(2) GT_CASE(3) GT_CASE(4) GT_CASE(5) GT_CASE(6) GT_CASE(7) GT_CASE(8) GT_CASE(9)
#undef GT_CASE
    throw gt_error(GT_ERR_UNSUPPORTED, "unsupported (D, window) combination");
}

template <template <int> class L, class... A>
void dispatch_fast(int KW, A&&... a) {
    switch (KW) {
    case 3: L<3>::run(std::forward<A>(a)...); return;
    case 4: L<4>::run(std::forward<A>(a)...); return;
    case 5: L<5>::run(std::forward<A>(a)...); return;
    case 6: L<6>::run(std::forward<A>(a)...); return;
    case 7: L<7>::run(std::forward<A>(a)...); return;
    default: throw gt_error(GT_ERR_UNSUPPORTED, "unsupported window for the 2D hot path");
    }
}

constexpr int kC2NCChunk = 8;

template <int D, int KW>
struct LaunchC2NC {
    static void run(const float2* grid, float2* out, const float* pos, int M, int F, int C,
                    const gt::GridGeom& g, int sign, const float* dcw, long long dcw_len,
                    int accumulate, hipStream_t s) {
        const long long n = (long long)M * F;
        hipLaunchKernelGGL((gt::c2nc_gather_kernel<D, KW, kC2NCChunk>), dim3((unsigned)((n + 255) / 256)),
                           dim3(256), 0, s, grid, out, pos, M, F, C, g, sign, dcw, dcw_len, accumulate);
    }
};

template <int D, int KW>
struct LaunchTrajGrad {
    static void run(const float2* grid, const float2* gy, const float* pos, int M, int F, int C,
                    const gt::GridGeom& g, const float* dcw, long long dcw_len, float* out, hipStream_t s) {
        const long long n = (long long)M * F;
        hipLaunchKernelGGL((gt::traj_grad_kernel<D, KW>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, grid,
                           gy, pos, M, F, C, g, 1, dcw, dcw_len, out);
    }
};

template <int D, int KW>
struct LaunchNC2CTile {
    static void run(const float2* samples, float2* grid, float2* partials, const float* pos,
                    const uint32_t* tile_samples, const gt::WorkItem* items, int n_items, int M,
                    int F, int C, int CC, const gt::GridGeom& g, int sign, const float* dcw,
                    long long dcw_len, int accumulate, hipStream_t s) {
        const unsigned chunks = (unsigned)((C + CC - 1) / CC);
        const size_t lds = size_t(2) * CC * g.tile_cells * sizeof(float);
        hipLaunchKernelGGL((gt::nc2c_tile_kernel<D, KW>), dim3((unsigned)n_items, chunks), dim3(64), lds, s,
                           samples, grid, partials, pos, tile_samples, items, M, F, C, CC, g, sign,
                           dcw, dcw_len, accumulate);
    }
};

template <int D, int KW>
struct LaunchNC2CAtomic {
    static void run(const float2* samples, float2* grid, const float* pos, int M, int F, int C,
                    const gt::GridGeom& g, int sign, const float* dcw, long long dcw_len,
                    hipStream_t s) {
        const long long n = (long long)M * F;
        hipLaunchKernelGGL((gt::nc2c_atomic_kernel<D, KW>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                           s, samples, grid, pos, M, F, C, g, sign, dcw, dcw_len);
    }
};

template <int D, int KW>
struct LaunchC2NCD {
    static void run(const double2* grid, double2* out, const double* pos, int M, int F, int C,
                    const gt::GridGeom& g, int sign, const double* dcw, long long dcw_len, int accumulate,
                    hipStream_t s) {
        const long long n = (long long)M * F;
        hipLaunchKernelGGL((gt::c2nc_gather_d_kernel<D, KW, 4>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                           s, grid, out, pos, M, F, C, g, sign, dcw, dcw_len, accumulate);
    }
};
template <int D, int KW>
struct LaunchNC2CTileD {
    static void run(const double2* samples, double2* grid, double2* partials, const double* pos,
                    const uint32_t* tile_samples, const gt::WorkItem* items, int n_items, int M, int F, int C,
                    int CC, const gt::GridGeom& g, int sign, const double* dcw, long long dcw_len, int accumulate,
                    hipStream_t s) {
        const unsigned chunks = (unsigned)((C + CC - 1) / CC);
        const size_t lds = size_t(2) * CC * g.tile_cells * sizeof(double);
        hipLaunchKernelGGL((gt::nc2c_tile_d_kernel<D, KW>), dim3((unsigned)n_items, chunks), dim3(64), lds, s,
                           samples, grid, partials, pos, tile_samples, items, M, F, C, CC, g, sign, dcw, dcw_len,
                           accumulate);
    }
};

template <int KW>
struct LaunchTablesNC2C {
    static void run(const float* pos, const uint32_t* list, const uint32_t* keys, long long n,
                    const gt::GridGeom& g, float* wt, uint32_t* hd, float csc, hipStream_t s) {
        hipLaunchKernelGGL((gt::nc2c_table2d_kernel<KW>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s,
                           pos, list, keys, n, g, wt, hd, csc);
    }
};

template <int KW>
struct LaunchTablesC2NC {
    static void run(const float* pos, const uint32_t* list, long long n, const gt::GridGeom& g, uint32_t* tab,
                    hipStream_t s) {
        hipLaunchKernelGGL((gt::c2nc_table2d_kernel<KW>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s,
                           pos, list, n, g, tab);
    }
};

// The 2D gridding of many coils (NC2C from 9, C2NC from 17) runs on the
// matrix cores unless GT_NUFFT_MFMA=0 (the FP32 kernels, kept for comparison).
bool gridding_on_matrix_cores() {
    static const bool on = [] {
        const char* e = std::getenv("GT_NUFFT_MFMA");
        return !(e && e[0] == '0');
    }();
    return on;
}

// Coils per wave of the matrix-core NC2C (GT_NUFFT_MFMA_CW=16 or 32).
int mfma_coils_per_wave() {
    static const int cw = [] {
        const char* e = std::getenv("GT_NUFFT_MFMA_CW");
        return e && std::atoi(e) == 32 ? 32 : 16;
    }();
    return cw;
}

// Power of two that brings the largest 1-D kernel weight along axis d (at
// u = 0) to at most 1.
float kb_axis_scale(const gt_nufft_plan* p, int d) {
    const gt::GridGeom& g = p->geom;
    const double peak = double(gt::kb1(0.0f, g.nos_f[d], g.one_over_W, g.beta[d]));
    return std::ldexp(1.0f, -int(std::ceil(std::log2(peak))));
}

// Power of two that brings the largest product of the two 1-D kernel weights
// (at u = 0) to at most 1: the weight normaliser of nc2c_mfma2d_kernel.
float mfma_weight_scale(const gt_nufft_plan* p) {
    const gt::GridGeom& g = p->geom;
    const double peak = double(gt::kb1(0.0f, g.nos_f[0], g.one_over_W, g.beta[0])) *
                        double(gt::kb1(0.0f, g.nos_f[1], g.one_over_W, g.beta[1]));
    return std::ldexp(1.0f, -int(std::ceil(std::log2(peak))));
}

// Coil groups (NC2C) / sample slots (C2NC) of the 2D gridding kernels: a wave
// covers 32 / G coils, so fewer coils leave no lanes idle.
int coil_groups(int C) { return C <= 8 ? 4 : (C <= 16 ? 2 : 1); }

template <int KW>
struct LaunchWTableC2NC {
    static void run(const float* pos, const uint32_t* list, long long n, const gt::GridGeom& g, float sx, float sy,
                    float* wtab, hipStream_t s) {
        hipLaunchKernelGGL((gt::c2nc_wtable2d_kernel<KW>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, pos,
                           list, n, g, sx, sy, wtab);
    }
};

template <int KW>
struct LaunchFastNC2C {
    template <int G>
    static auto pick(const float* dcw, long long dcw_len, long long FM) {
        return !dcw ? gt::nc2c_reg2d_kernel<0, G>
                    : (dcw_len == FM ? gt::nc2c_reg2d_kernel<1, G> : gt::nc2c_reg2d_kernel<2, G>);
    }
    static void run(const float* samples, float* grid, float* partials, const gt_nufft_plan* p, int C,
                    const gt::Layout& L, int sign, const float* dcw, long long dcw_len, int accumulate,
                    hipStream_t s) {
        const int G = coil_groups(C);
        const int cw = gt::kCoilsPerWave / G;
        const unsigned chunks = (unsigned)((C + cw - 1) / cw);
        const long long FM = (long long)(p->M * p->F);
        // 9 coils or more fill the 16-coil matrix-core kernel's rows at least half way
        if ((G == 1 || (G == 2 && mfma_coils_per_wave() == 16)) && gridding_on_matrix_cores()) {
            // the tile product on the matrix cores, CW coils per wave
            const int cwm = mfma_coils_per_wave();
            const unsigned mchunks = (unsigned)((C + cwm - 1) / cwm);
            auto km = cwm == 16 ? (!dcw ? gt::nc2c_mfma2d_kernel<0, 16>
                                        : (dcw_len == FM ? gt::nc2c_mfma2d_kernel<1, 16> : gt::nc2c_mfma2d_kernel<2, 16>))
                                : (!dcw ? gt::nc2c_mfma2d_kernel<0, 32>
                                        : (dcw_len == FM ? gt::nc2c_mfma2d_kernel<1, 32> : gt::nc2c_mfma2d_kernel<2, 32>));
            const unsigned nblk = (unsigned)((p->bn.n_items + 7) / 8 * 8) * mchunks;  // items x chunks, see the kernel
            hipLaunchKernelGGL(km, dim3(nblk), dim3(64), 0, s,
                               samples, grid, partials, p->bn.table.as<float>(), p->bn.heads.as<uint32_t>(),
                               p->bn.items.as<gt::WorkItem>(), int(p->M), int(p->F), C, p->bn.geom, L, sign, dcw,
                               dcw_len, accumulate, mfma_weight_scale(p), p->bn.n_items, int(mchunks));
            return;
        }
        auto k = G == 4 ? pick<4>(dcw, dcw_len, FM) : (G == 2 ? pick<2>(dcw, dcw_len, FM) : pick<1>(dcw, dcw_len, FM));
        hipLaunchKernelGGL(k, dim3((unsigned)p->bn.n_items, chunks), dim3(64), 0, s,
                           samples, grid, partials, p->bn.table.as<float>(), p->bn.heads.as<uint32_t>(),
                           p->bn.items.as<gt::WorkItem>(),
                           int(p->M), int(p->F), C, p->bn.geom, L, sign, dcw, dcw_len, accumulate,
                           1.0f / mfma_weight_scale(p));
    }
};

template <int KW>
struct LaunchFastC2NC {
    template <int G>
    static auto pick(int mode) {
        return mode == 0 ? gt::c2nc_multi2d_kernel<KW, 0, G>
                         : (mode == 1 ? gt::c2nc_multi2d_kernel<KW, 1, G> : gt::c2nc_multi2d_kernel<KW, 2, G>);
    }
    static void run(const float* grid, float* out, gt_nufft_plan* p, int C, const gt::Layout& L, int sign,
                    const float* dcw, long long dcw_len, int accumulate, hipStream_t s) {
        Binning& b = p->bc;
        const int G = coil_groups(C);
        const int cw = gt::kCoilsPerWave / G;
        const unsigned chunks = (unsigned)((C + cw - 1) / cw);
        if (G == 1 && p->bm.wtab.ptr && gridding_on_matrix_cores()) {
            Binning& b = p->bm;
            // 17 coils or more: the home tiles' products on the matrix cores,
            // cwm coils per workgroup (9 to 16 coils measured faster on the
            // FP32 gather: 20.6 against 22.7 us at 256^2, 16 coils)
            const float oscale = 1.0f / (kb_axis_scale(p, 0) * kb_axis_scale(p, 1));  // powers of two
            const int cwm = mfma_coils_per_wave();
            const unsigned mchunks = (unsigned)((C + cwm - 1) / cwm);
            auto km = cwm == 16 ? gt::c2nc_mfma2d_kernel<KW, 16> : gt::c2nc_mfma2d_kernel<KW, 32>;
            const unsigned nblk = (unsigned)((b.n_items + 7) / 8 * 8) * mchunks;  // items x chunks, see the kernel
            hipLaunchKernelGGL(km, dim3(nblk), dim3(256), size_t(gt::c2nc_mfma_lds_bytes(KW, cwm)), s, grid, out,
                               b.wtab.as<float>(), b.list.as<uint32_t>(), b.items.as<gt::WorkItem>(), int(p->M),
                               int(p->F), C, b.geom, L, sign, dcw, dcw_len, accumulate, oscale, b.n_items,
                               int(mchunks));
            return;
        }
        const size_t lds = size_t(gt::c2nc_lds_bytes(KW, G));
        const int mode = accumulate ? 2 : (!dcw ? 0 : (dcw_len == (long long)(p->M * p->F) ? 1 : 2));
        // a per-sample dcw is gathered per table entry inside the kernel
        auto k = G == 4 ? pick<4>(mode) : (G == 2 ? pick<2>(mode) : pick<1>(mode));
        hipLaunchKernelGGL(k, dim3((unsigned)b.n_items, chunks), dim3(64 * gt::kWaves), lds, s, grid, out,
                           b.table.as<uint32_t>(), b.items.as<gt::WorkItem>(), int(p->M), int(p->F), C, b.geom, L,
                           sign, dcw, dcw_len, b.dcw_e.as<float>(), accumulate);
    }
};

struct LaunchTablesNC2C3 {
    static void run(const float* pos, const uint32_t* list, const uint32_t* keys, long long n,
                    const gt::GridGeom& g, float* wt, uint32_t* hd, hipStream_t s) {
        hipLaunchKernelGGL(gt::nc2c_table3d_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, pos, list,
                           keys, n, g, wt, hd);
    }
};

template <int KW>
struct LaunchWTableC2NC3 {
    static void run(const float* pos, const uint32_t* list, long long n, const gt::GridGeom& g, float sx, float sy,
                    float* wtab, int ey, hipStream_t s) {
        const dim3 blocks((unsigned)((n + 255) / 256));
        if (ey == 16)
            hipLaunchKernelGGL((gt::c2nc_wtable3d_kernel<KW, 16>), blocks, dim3(256), 0, s, pos, list, n, g, sx, sy, wtab);
        else if (ey == 12)
            hipLaunchKernelGGL((gt::c2nc_wtable3d_kernel<KW, 12>), blocks, dim3(256), 0, s, pos, list, n, g, sx, sy, wtab);
        else
            hipLaunchKernelGGL((gt::c2nc_wtable3d_kernel<KW, 8>), blocks, dim3(256), 0, s, pos, list, n, g, sx, sy, wtab);
    }
};

// Staged rows in y of the matrix-core 3D C2NC: 12 (tiles of 3 x 7 x 3 cells
// at windows of 5 or 6 cells); GT_NUFFT_C2NC3_EY = 8 or 16 selects the others.
int c2nc3_rows() {
    static const int ey = [] {
        const char* e = std::getenv("GT_NUFFT_C2NC3_EY");
        const int v = e ? std::atoi(e) : 12;
        return (v == 8 || v == 16) ? v : 12;
    }();
    return ey;
}

int c2nc3_table_floats(int ey) { return 2 * gt::k3ME + ey; }

template <int KW>
struct LaunchTablesC2NC3 {
    static void run(const float* pos, const uint32_t* list, long long n, const gt::GridGeom& g, uint32_t* tab,
                    hipStream_t s) {
        hipLaunchKernelGGL((gt::c2nc_table3d_kernel<KW>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s,
                           pos, list, n, g, tab);
    }
};

// Power of two that brings the largest product of the three 1-D kernel
// weights (at u = 0) to at most 1: the weight normaliser of nc2c_mfma3d_kernel.
float mfma_weight_scale3(const gt_nufft_plan* p) {
    const gt::GridGeom& g = p->geom;
    double peak = 1.0;
    for (int d = 0; d < 3; d++) peak *= double(gt::kb1(0.0f, g.nos_f[d], g.one_over_W, g.beta[d]));
    return std::ldexp(1.0f, -int(std::ceil(std::log2(peak))));
}

void launch_fast3_nc2c(const float* samples, float* grid, float* partials, const gt_nufft_plan* p, int C,
                       const gt::Layout& L, int sign, const float* dcw, long long dcw_len, int accumulate,
                       hipStream_t s) {
    const unsigned chunks = (unsigned)((C + gt::k3Coils - 1) / gt::k3Coils);
    const long long FM = (long long)(p->M * p->F);
    if (gridding_on_matrix_cores()) {
        // the tile products on the matrix cores (split fp16, FP32 sums)
        auto km = !dcw ? gt::nc2c_mfma3d_kernel<0>
                       : (dcw_len == FM ? gt::nc2c_mfma3d_kernel<1> : gt::nc2c_mfma3d_kernel<2>);
        hipLaunchKernelGGL(km, dim3((unsigned)p->bn.n_items, chunks), dim3(64), 0, s, samples, grid, partials,
                           p->bn.table.as<float>(), p->bn.heads.as<uint32_t>(), p->bn.items.as<gt::WorkItem>(),
                           int(p->M), int(p->F), C, p->bn.geom, L, sign, dcw, dcw_len, accumulate,
                           mfma_weight_scale3(p));
        return;
    }
    auto k = !dcw ? gt::nc2c_reg3d_kernel<0> : (dcw_len == FM ? gt::nc2c_reg3d_kernel<1> : gt::nc2c_reg3d_kernel<2>);
    hipLaunchKernelGGL(k, dim3((unsigned)p->bn.n_items, chunks), dim3(64), 0, s, samples, grid, partials,
                       p->bn.table.as<float>(), p->bn.heads.as<uint32_t>(), p->bn.items.as<gt::WorkItem>(),
                       int(p->M), int(p->F), C, p->bn.geom, L, sign, dcw, dcw_len, accumulate);
}

template <int KW>
struct LaunchFast3C2NC {
    static void run(const float* grid, float* out, gt_nufft_plan* p, int C, const gt::Layout& L, int sign,
                    const float* dcw, long long dcw_len, int accumulate, hipStream_t s) {
        const unsigned chunks = (unsigned)((C + gt::k3Coils - 1) / gt::k3Coils);
        const size_t lds = size_t(gt::c2nc3_lds_bytes(KW));
        const int mode = accumulate ? 2 : (!dcw ? 0 : (dcw_len == (long long)(p->M * p->F) ? 1 : 2));
        Binning& b = p->bc;
        if (p->c2nc3_mfma) {
            // persistent: as many workgroups as are resident at once (LDS and
            // registers), a multiple of the 8 XCDs so that every item stays on
            // the XCD its position assigns
            const float oscale = 1.0f / (kb_axis_scale(p, 0) * kb_axis_scale(p, 1));  // powers of two
            auto pick = [mode](auto k0, auto k1, auto k2) { return mode == 0 ? k0 : (mode == 1 ? k1 : k2); };
            auto km = p->c2nc3_ey == 16
                          ? pick(gt::c2nc_mfma3d_kernel<0, 16>, gt::c2nc_mfma3d_kernel<1, 16>,
                                 gt::c2nc_mfma3d_kernel<2, 16>)
                          : (p->c2nc3_ey == 12 ? pick(gt::c2nc_mfma3d_kernel<0, 12>, gt::c2nc_mfma3d_kernel<1, 12>,
                                                      gt::c2nc_mfma3d_kernel<2, 12>)
                                               : pick(gt::c2nc_mfma3d_kernel<0, 8>, gt::c2nc_mfma3d_kernel<1, 8>,
                                                      gt::c2nc_mfma3d_kernel<2, 8>));
            int per_cu = 0;
            GT_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, km, 64 * gt::k3MWaves, 0));
            const int resident = std::max(8, (std::max(per_cu, 1) * p->cus) & ~7);
            const unsigned ctas = (unsigned)std::min(((b.n_items + 7) / 8) * 8, resident);
            if (mode == 1) {
                b.dcw_e.reserve(std::max<size_t>(b.n, 1) * sizeof(float));
                hipLaunchKernelGGL(gt::entry_dcw_kernel, dim3((unsigned)((b.n + 255) / 256)), dim3(256), 0, s,
                                   b.list.as<uint32_t>(), (long long)b.n, dcw, b.dcw_e.as<float>());
            }
            hipLaunchKernelGGL(km, dim3(ctas, chunks), dim3(64 * gt::k3MWaves), 0, s, grid, out, b.wtab.as<float>(),
                               b.list.as<uint32_t>(), b.recs.as<uint32_t>(), b.n_items, int(p->M), int(p->F), C,
                               b.geom, L, sign, dcw, dcw_len, b.dcw_e.as<float>(), accumulate, oscale);
            return;
        }
        if (mode == 1) {
            b.dcw_e.reserve(std::max<size_t>(b.n, 1) * sizeof(float));
            hipLaunchKernelGGL(gt::entry_dcw_kernel, dim3((unsigned)((b.n + 255) / 256)), dim3(256), 0, s,
                               b.list.as<uint32_t>(), (long long)b.n, dcw, b.dcw_e.as<float>());
        }
        // persistent: two workgroups per CU (their LDS), a multiple of the 8
        // XCDs so that every item stays on the XCD its position assigns
        const int resident = std::max(8, (2 * p->cus) & ~7);
        const unsigned ctas = (unsigned)std::min(b.n_items, resident);
        auto k = mode == 0 ? gt::c2nc_multi3d_kernel<KW, 0>
                           : (mode == 1 ? gt::c2nc_multi3d_kernel<KW, 1> : gt::c2nc_multi3d_kernel<KW, 2>);
        hipLaunchKernelGGL(k, dim3(ctas, chunks), dim3(64 * gt::k3Waves), lds, s, grid, out, b.table.as<uint32_t>(),
                           b.items.as<gt::WorkItem>(), b.n_items, int(p->M), int(p->F), C, b.geom, L, sign, dcw,
                           dcw_len, b.dcw_e.as<float>(), accumulate);
    }
};

int coil_chunk(const gt_nufft_plan* p) {
    // LDS budget per one-wave workgroup: 16 KiB of re/im accumulators
    int cc = int(16384 / (8 * std::max(1, p->geom.tile_cells)));
    return std::max(1, std::min(cc, 32));
}

gt::GridGeom make_geom(const gt_nufft_plan* p, const int* tile, bool tile_covers_window = true) {
    gt::GridGeom g{};
    g.D = p->D;
    g.radius = p->radius;
    g.one_over_W = p->one_over_W;
    g.radius_d = p->radius_d;
    g.one_over_W_d = p->one_over_W_d;
    g.G = (long long)p->G();
    g.ntiles_total = 1;
    g.tile_cells = 1;
    for (int d = 0; d < 4; d++) {
        if (d < p->D) {
            g.nos[d] = int(p->Nos[d]);
            g.nos_f[d] = float(unsigned(p->Nos[d]));
            g.beta[d] = p->beta[d];
            g.nos_d[d] = double(p->Nos[d]);
            g.beta_d[d] = p->beta_d[d];
            // the generic kernels: a window of KW cells touches at most two tiles
            // per dimension; the 3D hot path's tables take any tile
            const int t = std::min<int>(tile_covers_window ? std::max(tile[d], p->KW) : tile[d], int(p->Nos[d]));
            g.tile[d] = t;
            g.ntiles[d] = int((p->Nos[d] + t - 1) / t);
        } else {
            g.nos[d] = 1;
            g.nos_f[d] = 1.0f;
            g.nos_d[d] = 1.0;
            g.tile[d] = 1;
            g.ntiles[d] = 1;
        }
        g.ntiles_total *= g.ntiles[d];
        g.tile_cells *= g.tile[d];
    }
    return g;
}

void setup_geometry(gt_nufft_plan* p) {
    int tile[4];
    for (int d = 0; d < 4; d++) {
        // D = 4: tiles as wide as the window (a tile of KW^4 cells per wavefront)
        const int def = p->D == 1 ? 256 : (p->D == 2 ? 16 : (p->D == 3 ? 8 : 1));
        tile[d] = p->tile_req[d] > 0 ? p->tile_req[d] : def;
    }
    p->geom = make_geom(p, tile);
    // the generic NC2C holds at least one coil's tile of re/im accumulators in LDS
    GT_REQUIRE(size_t(2) * (p->dbl ? sizeof(double) : sizeof(float)) * size_t(p->geom.tile_cells) <= 65536,
               GT_ERR_UNSUPPORTED, "the gridding tile does not fit in LDS (4-D double plans take windows of <= 8 cells)");
    // 2D hot path: STANDARD gridding, a window of 3..7 cells, 8 x 8 tiles
    // (the default; any other requested tile selects the generic kernels) and
    // a grid of at least tile + window cells per dimension.
    p->fast2d = false;
    if (p->D == 2 && !p->dbl && p->conv_type != GT_CONV_ATOMIC && p->KW >= 3 && p->KW <= gt::kTabKW) {
        int ft[3];
        for (int d = 0; d < 3; d++) ft[d] = p->tile_req[d] > 0 ? p->tile_req[d] : gt::kRT;
        bool ok = true;
        for (int d = 0; d < 2; d++) {
            ok &= ft[d] == gt::kRT;
            ok &= int(p->Nos[d]) >= ft[d] + p->KW;
        }
        ok &= gt::c2nc_lds_bytes(p->KW) <= 65536;
        // window origins are found with ceil(p - r) in float: keep r clear of
        // an integer by more than the rounding of p (or exactly on one)
        const double fr = double(p->radius) - std::floor(double(p->radius));
        const double eps = 4.0 * std::ldexp(1.0, -23) * double(std::max(p->Nos[0], p->Nos[1]));
        ok &= fr == 0.0 || (fr > eps && fr < 1.0 - eps);
        if (ok) {
            p->fast2d = true;
            p->bn.geom = make_geom(p, ft);
            p->bc.geom = make_geom(p, ft);
            // the matrix-core C2NC stages rows of 16 cells: home tiles of
            // 16 - c2nc_extent(0, KW) columns (11 at windows of 5 or 6 cells)
            // fill them; 8 rows as the FP32 kernel's.  Its staged columns run
            // to tile origin - floor(r) + 15, within one wrap of the grid for
            // N_os >= 32 (smaller grids keep 8-column tiles)
            const int fm[3] = {int(p->Nos[0]) >= 32 ? gt::c2nc_mfma_tile_x(p->KW) : gt::kRT, gt::kRT, 1};
            p->bm.geom = make_geom(p, fm);
        }
    }
    // 3D hot path: STANDARD gridding, a window of 3..7 cells, no requested
    // tile, and a grid of at least tile + window cells per dimension.
    p->fast3d = false;
    p->c2nc3_mfma = false;
    if (p->D == 3 && !p->dbl && p->conv_type != GT_CONV_ATOMIC && p->KW >= 3 && p->KW <= gt::kTabKW &&
        p->tile_req[0] <= 0 && p->tile_req[1] <= 0 && p->tile_req[2] <= 0) {
        const int tn[3] = {gt::k3T0, gt::k3T1, gt::k3T2}, tc[3] = {gt::k3S0, gt::k3S1, gt::k3S2};
        bool ok = true;
        for (int d = 0; d < 3; d++) ok &= int(p->Nos[d]) >= std::max(tn[d], tc[d]) + p->KW;
        const double fr = double(p->radius) - std::floor(double(p->radius));
        const double eps = 4.0 * std::ldexp(1.0, -23) * double(std::max({p->Nos[0], p->Nos[1], p->Nos[2]}));
        ok &= fr == 0.0 || (fr > eps && fr < 1.0 - eps);
        ok &= gt::c2nc3_lds_bytes(p->KW) <= 65536;
        if (ok) {
            p->fast3d = true;
            p->bn.geom = make_geom(p, tn, false);
            // the matrix-core C2NC stages tile + halo as 8 cells per dimension:
            // tiles of 3 cells for windows of 5 or 6 cells, 5 for 3 or 4
            const int tm = gt::c2nc_mfma3_tile(p->KW);
            // taller tiles in y where the grid holds their staged rows within one wrap
            p->c2nc3_ey = int(p->Nos[1]) >= c2nc3_rows() + p->KW ? c2nc3_rows() : 8;
            const int tmc[3] = {tm, gt::c2nc_mfma3_tile(p->KW, p->c2nc3_ey), tm};
            p->c2nc3_mfma = gridding_on_matrix_cores() && tm >= 3;
            p->bc.geom = make_geom(p, p->c2nc3_mfma ? tmc : tc, false);
            if (p->c2nc3_mfma) {
                // keys in column order (columns in bricks of 4 x 4): the key space is padded to whole bricks
                gt::GridGeom& g = p->bc.geom;
                g.ntiles_total = 16 * ((g.ntiles[0] + 3) / 4) * ((g.ntiles[1] + 3) / 4) * g.ntiles[2];
            }
        }
    }
}

// ---- preprocess ------------------------------------------------------------
// Sorts (key, sample) pairs by key (stable, so every tile keeps ascending
// sample order) and returns the per-key counts on the host.
std::vector<uint32_t> sort_pairs(DevBuf& keys, DevBuf& vals, size_t npairs, uint32_t nkeys, DevBuf& keys_out,
                                 DevBuf& vals_out, hipStream_t s) {
    DevBuf temp, hist;
    keys_out.reserve(std::max<size_t>(npairs, 1) * 4);
    vals_out.reserve(std::max<size_t>(npairs, 1) * 4);
    int end_bit = 1;
    while ((uint64_t(1) << end_bit) < uint64_t(nkeys)) end_bit++;
    size_t tb = 0;
    GT_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, keys.as<uint32_t>(), keys_out.as<uint32_t>(),
                                              vals.as<uint32_t>(), vals_out.as<uint32_t>(), (int)npairs, 0,
                                              end_bit, s));
    temp.reserve(std::max<size_t>(tb, 1));
    GT_HIP(hipcub::DeviceRadixSort::SortPairs(temp.ptr, tb, keys.as<uint32_t>(), keys_out.as<uint32_t>(),
                                              vals.as<uint32_t>(), vals_out.as<uint32_t>(), (int)npairs, 0,
                                              end_bit, s));
    hist.reserve(size_t(nkeys) * 4 * 3);  // counts, first, last
    uint32_t* cnt = hist.as<uint32_t>();
    GT_HIP(hipMemsetAsync(cnt + nkeys, 0, size_t(nkeys) * 4 * 2, s));
    if (npairs)
        hipLaunchKernelGGL(gt::run_bounds_kernel, dim3((unsigned)((npairs + 255) / 256)), dim3(256), 0, s,
                           keys_out.as<uint32_t>(), (long long)npairs, cnt + nkeys, cnt + 2 * size_t(nkeys));
    hipLaunchKernelGGL(gt::run_counts_kernel, dim3((nkeys + 255) / 256), dim3(256), 0, s, cnt + nkeys,
                       cnt + 2 * size_t(nkeys), nkeys, cnt);
    GT_HIP(hipGetLastError());
    std::vector<uint32_t> counts(nkeys, 0);
    GT_HIP(hipMemcpyAsync(counts.data(), hist.ptr, size_t(nkeys) * 4, hipMemcpyDeviceToHost, s));
    GT_HIP(hipStreamSynchronize(s));
    temp.release();
    hist.release();
    return counts;
}

// Window-tile pairs: one (tile, sample) pair for every tile a sample's window touches.
std::vector<uint32_t> bin_windows(gt_nufft_plan* p, const gt::GridGeom& g, Binning& b, DevBuf* keys_keep,
                                  hipStream_t s) {
    const long long nsamp = (long long)(p->M * p->F);
    const uint32_t nkeys = uint32_t(p->F) * uint32_t(g.ntiles_total);
    DevBuf counts, offsets, keys, vals, keys_sorted, temp;
    counts.reserve(nsamp * sizeof(uint32_t));
    offsets.reserve(nsamp * sizeof(uint32_t));
    hipLaunchKernelGGL(gt::count_tiles_kernel, dim3((unsigned)((nsamp + 255) / 256)), dim3(256), 0, s,
                       p->traj.as<float>(), nsamp, g, counts.as<uint32_t>());
    GT_HIP(hipGetLastError());
    size_t tb = 0;
    GT_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, counts.as<uint32_t>(), offsets.as<uint32_t>(),
                                            (int)nsamp, s));
    temp.reserve(std::max<size_t>(tb, 1));
    GT_HIP(hipcub::DeviceScan::ExclusiveSum(temp.ptr, tb, counts.as<uint32_t>(), offsets.as<uint32_t>(),
                                            (int)nsamp, s));
    uint32_t last_off = 0, last_cnt = 0;
    GT_HIP(hipMemcpyAsync(&last_off, offsets.as<uint32_t>() + nsamp - 1, 4, hipMemcpyDeviceToHost, s));
    GT_HIP(hipMemcpyAsync(&last_cnt, counts.as<uint32_t>() + nsamp - 1, 4, hipMemcpyDeviceToHost, s));
    GT_HIP(hipStreamSynchronize(s));
    const size_t npairs = size_t(last_off) + last_cnt;
    GT_REQUIRE(npairs < (size_t(1) << 31), GT_ERR_UNSUPPORTED, "too many sample-tile pairs");
    keys.reserve(std::max<size_t>(npairs, 1) * 4);
    vals.reserve(std::max<size_t>(npairs, 1) * 4);
    hipLaunchKernelGGL(gt::emit_pairs_kernel, dim3((unsigned)((nsamp + 255) / 256)), dim3(256), 0, s,
                       p->traj.as<float>(), nsamp, (long long)p->M, g, offsets.as<uint32_t>(),
                       keys.as<uint32_t>(), vals.as<uint32_t>());
    GT_HIP(hipGetLastError());
    DevBuf& ko = keys_keep ? *keys_keep : keys_sorted;
    auto cnt = sort_pairs(keys, vals, npairs, nkeys, ko, b.list, s);
    b.n = npairs;
    return cnt;
}

// Home-tile pairs: one pair per sample (2D hot path C2NC).
std::vector<uint32_t> bin_home(gt_nufft_plan* p, const gt::GridGeom& g, Binning& b, hipStream_t s) {
    const long long nsamp = (long long)(p->M * p->F);
    const uint32_t nkeys = uint32_t(p->F) * uint32_t(g.ntiles_total);
    DevBuf keys, vals, keys_sorted;
    keys.reserve(nsamp * 4);
    vals.reserve(nsamp * 4);
    if (p->D == 2)
        hipLaunchKernelGGL(gt::home_keys_kernel, dim3((unsigned)((nsamp + 255) / 256)), dim3(256), 0, s,
                           p->traj.as<float>(), nsamp, (long long)p->M, g, keys.as<uint32_t>(), vals.as<uint32_t>());
    else if (p->c2nc3_mfma)
        hipLaunchKernelGGL(gt::home_keys_column3_kernel, dim3((unsigned)((nsamp + 255) / 256)), dim3(256), 0, s,
                           p->traj.as<float>(), nsamp, (long long)p->M, g, keys.as<uint32_t>(), vals.as<uint32_t>());
    else
        hipLaunchKernelGGL(gt::home_keys_nd_kernel, dim3((unsigned)((nsamp + 255) / 256)), dim3(256), 0, s,
                           p->traj.as<float>(), nsamp, (long long)p->M, g, keys.as<uint32_t>(), vals.as<uint32_t>());
    GT_HIP(hipGetLastError());
    auto cnt = sort_pairs(keys, vals, size_t(nsamp), nkeys, keys_sorted, b.list, s);
    b.n = size_t(nsamp);
    return cnt;
}

// Workgroup order.  MI355X hands workgroup i of a launch to XCD i mod 8, and
// each XCD has its own L2.  Items (created in key order: frame, tile row, tile
// column) are cut into kXcds contiguous runs of equal work -- horizontal bands
// of tiles, so neighbouring tiles, which share halo cells and samples, meet in
// one L2 -- each run is sorted longest first (the dispatcher hands out
// workgroups in index order), and the runs are interleaved, padded with no-op
// items so that item i always runs on XCD i mod kXcds.  GT_NUFFT_XCD=0 keeps
// a single longest-first order instead.  The 3D C2NC keeps key order within a
// run (no longest-first sort): its items are short, and neighbouring tiles
// then stage overlapping halos while those are in the XCD's L2.
constexpr int kXcds = 8;

bool xcd_ordering() {
    static const bool on = [] {
        const char* e = std::getenv("GT_NUFFT_XCD");
        return !(e && std::string(e) == "0");
    }();
    return on;
}

std::vector<gt::WorkItem> xcd_order(std::vector<gt::WorkItem> items, bool sort_longest) {
    auto longest_first = [sort_longest](std::vector<gt::WorkItem>& v) {
        if (!sort_longest) return;
        std::stable_sort(v.begin(), v.end(), [](const gt::WorkItem& a, const gt::WorkItem& b) {
            return (a.end - a.begin) > (b.end - b.begin);
        });
    };
    if (!xcd_ordering() || items.size() < size_t(4 * kXcds)) {
        longest_first(items);
        return items;
    }
    // a fixed cost per item (staging, write-out) plus one unit per entry
    auto work = [](const gt::WorkItem& w) { return double(w.end - w.begin) + 32.0; };
    double total = 0.0;
    for (const auto& w : items) total += work(w);
    std::vector<std::vector<gt::WorkItem>> run(kXcds);
    double cum = 0.0;
    for (const auto& w : items) {
        const int g = std::min(kXcds - 1, int(cum / total * kXcds));
        run[size_t(g)].push_back(w);
        cum += work(w);
    }
    size_t longest = 0;
    for (auto& r : run) {
        longest_first(r);
        longest = std::max(longest, r.size());
    }
    std::vector<gt::WorkItem> out;
    out.reserve(longest * kXcds);
    for (size_t k = 0; k < longest; k++)
        for (int g = 0; g < kXcds; g++)
            out.push_back(k < run[size_t(g)].size() ? run[size_t(g)][k] : gt::WorkItem{0, 0, 0, gt::kNoopItem});
    return out;
}

// Work items over per-key counts: one per key, keys above L entries split into
// segments (with partial tiles summed in segment order when allowed).
void make_items(gt_nufft_plan* p, Binning& b, const std::vector<uint32_t>& counts, size_t L, bool partials,
                hipStream_t s) {
    std::vector<gt::WorkItem> items;
    std::vector<gt::RedItem> red;
    items.reserve(counts.size());
    uint32_t begin = 0;
    int n_partials = 0;
    for (uint32_t k = 0; k < uint32_t(counts.size()); k++) {
        const uint32_t n = counts[k];
        if (n == 0 && !partials) continue;  // C2NC: nothing to output for an empty tile
        if (n <= L) {
            items.push_back({k, begin, begin + n, -1});
        } else if (!partials) {
            for (uint32_t b0 = 0; b0 < n; b0 += uint32_t(L))
                items.push_back({k, begin + b0, begin + std::min<uint32_t>(n, b0 + uint32_t(L)), -1});
        } else {
            const uint32_t nseg = uint32_t((n + L - 1) / L);
            red.push_back({k, n_partials, int32_t(nseg), 0});
            for (uint32_t sgm = 0; sgm < nseg; sgm++) {
                const uint32_t b0 = begin + uint32_t(sgm * L);
                const uint32_t e0 = std::min<uint32_t>(begin + n, uint32_t(b0 + L));
                items.push_back({k, b0, e0, n_partials++});
            }
        }
        begin += n;
    }
    // the 3D C2NC items (short, staging a halo each) keep key order
    items = xcd_order(std::move(items), !(p->fast3d && !partials));
    if (items.empty()) items.push_back({0, 0, 0, gt::kNoopItem});  // keep one launchable item
    b.n_items = int(items.size());
    b.n_red = int(red.size());
    b.n_partials = n_partials;
    b.items.reserve(items.size() * sizeof(gt::WorkItem));
    b.red.reserve(std::max<size_t>(red.size(), 1) * sizeof(gt::RedItem));
    GT_HIP(hipMemcpyAsync(b.items.ptr, items.data(), items.size() * sizeof(gt::WorkItem), hipMemcpyHostToDevice,
                          s));
    if (!red.empty())
        GT_HIP(hipMemcpyAsync(b.red.ptr, red.data(), red.size() * sizeof(gt::RedItem), hipMemcpyHostToDevice, s));
    GT_HIP(hipStreamSynchronize(s));
}

// Work items of the matrix-core 3D C2NC: runs of consecutive non-empty tiles
// of one column segment -- the column's tiles in one z slab of kSlab tiles --
// holding at most L entries and k3MMaxTiles tiles; a tile above L entries
// becomes single-tile items of L entries.  Items run slab by slab, the
// columns of a slab in key order (bricks of 4 x 4 columns), so the columns
// staged at one time on an XCD are x and y neighbours at the same z and share
// their halo cells in its L2.  Items are XCD-ordered in that order and stored
// as records of k3MRec dwords (key, tiles, tile entry bounds).
void make_column_items(Binning& b, const std::vector<uint32_t>& counts, size_t L, hipStream_t s) {
    constexpr uint32_t kSlab = 8;
    const uint32_t ntz = uint32_t(b.geom.ntiles[2]);
    const uint32_t per_frame = uint32_t(b.geom.ntiles_total), ncol = per_frame / ntz;
    const uint32_t nf = uint32_t(counts.size() / per_frame);
    std::vector<uint32_t> tb(counts.size() + 1, 0);
    for (size_t k = 0; k < counts.size(); k++) tb[k + 1] = tb[k] + counts[k];
    std::vector<gt::WorkItem> items;  // key, begin, end, number of tiles
    for (uint32_t f = 0; f < nf; f++)
        for (uint32_t z0 = 0; z0 < ntz; z0 += kSlab)
            for (uint32_t c = 0; c < ncol; c++) {
                const uint32_t kb = f * per_frame + c * ntz;
                uint32_t k = kb + z0;
                const uint32_t ke = kb + std::min(ntz, z0 + kSlab);
                while (k < ke) {
                    const uint32_t n = counts[k];
                    if (n == 0) {
                        k++;
                        continue;
                    }
                    if (n > L) {
                        for (uint32_t b0 = tb[k]; b0 < tb[k + 1]; b0 += uint32_t(L))
                            items.push_back({k, b0, std::min<uint32_t>(tb[k + 1], b0 + uint32_t(L)), 1});
                        k++;
                        continue;
                    }
                    // extend along the column while the tiles are non-empty and fit
                    uint32_t e = k + 1, tot = n;
                    while (e < ke && counts[e] > 0 && tot + counts[e] <= L && e - k < uint32_t(gt::k3MMaxTiles))
                        tot += counts[e++];
                    items.push_back({k, tb[k], tb[e], int32_t(e - k)});
                    k = e;
                }
            }
    items = xcd_order(std::move(items), false);
    std::vector<uint32_t> rec(std::max<size_t>(items.size(), 1) * gt::k3MRec, 0u);
    for (size_t i = 0; i < items.size(); i++) {
        const gt::WorkItem& w = items[i];
        if (w.partial == gt::kNoopItem) continue;  // 0 tiles
        uint32_t* r = rec.data() + i * gt::k3MRec;
        r[0] = w.key;
        r[1] = uint32_t(w.partial);
        r[2] = w.begin;
        for (int t = 1; t < w.partial; t++) r[2 + t] = tb[w.key + uint32_t(t)];
        r[2 + w.partial] = w.end;
    }
    b.n_items = int(std::max<size_t>(items.size(), 1));
    b.n_red = 0;
    b.n_partials = 0;
    b.items.release();
    b.recs.reserve(rec.size() * sizeof(uint32_t));
    GT_HIP(hipMemcpyAsync(b.recs.ptr, rec.data(), rec.size() * sizeof(uint32_t), hipMemcpyHostToDevice, s));
    GT_HIP(hipStreamSynchronize(s));
}

// Entries per work item.  A wavefront walks its entries serially, so the
// longest item bounds the kernel time; splitting an NC2C tile costs one
// partial tile write and read, splitting a C2NC tile one more halo staging.
size_t segment_length(const gt_nufft_plan* p, const std::vector<uint32_t>& counts, size_t npairs,
                      bool nc2c) {
    if (p->c2nc3_mfma && !nc2c) {
        // the matrix-core 3D C2NC: an item restarts its column with 8 staged
        // planes, so longer items stage less (128^3 kooshball, 8 coils: 256
        // entries 1.37 ms, 512 1.25, 1024 1.21, 2048 1.19), but at least 3
        // items per CU for small problems
        const size_t target = npairs / size_t(3 * std::max(p->cus, 1));
        size_t L = 256;
        while (L * 2 <= std::min<size_t>(size_t(gt::k3MSeg), target)) L *= 2;
        return p->max_items_req > 0 ? std::min(L, size_t(p->max_items_req)) : L;
    }
    if ((p->fast2d || p->fast3d) && !nc2c) {
        // the other C2NC kernels stage a work item's whole table in LDS
        const size_t cap = size_t(p->fast2d ? gt::kC2NCSeg : gt::k3Seg);
        return p->max_items_req > 0 ? std::min(cap, size_t(p->max_items_req)) : cap;
    }
    if (p->nc2c_len_req > 0) return size_t(p->nc2c_len_req);
    if (p->max_items_req > 0) return size_t(p->max_items_req);
    if (p->fast2d) {
        // the matrix-core NC2C spends less per entry: items up to 512 entries
        // (fewer partial tiles), but at least 2 items per CU for small problems
        // (128^2, one frame of 16384 samples: 64-entry items 15 us, 512-entry 40 us)
        const size_t cap = gridding_on_matrix_cores() ? 512 : 256;
        const size_t target = npairs / size_t(2 * std::max(p->cus, 1));
        size_t L = 64;
        while (L * 2 <= std::min(cap, target)) L *= 2;
        return L;
    }
    if (p->fast3d) {
        // the matrix-core 3D NC2C: longer items mean fewer partial tiles to
        // write and sum (128^3 kooshball, 8 coils: 1024 entries 1.55 ms, 2048
        // 1.39, 4096 1.33, 8192 1.31), at least 2 items per CU for small problems
        if (!gridding_on_matrix_cores()) return 1024;
        const size_t target = npairs / size_t(2 * std::max(p->cus, 1));
        size_t L = 256;
        while (L * 2 <= std::min<size_t>(4096, target)) L *= 2;
        return L;
    }
    const double avg = double(npairs) / double(std::max<size_t>(counts.size(), 1));
    return size_t(std::max(256.0, 4.0 * avg));
}

void build_nc2c(gt_nufft_plan* p, hipStream_t s) {
    Binning& b = p->bn;
    if (p->fast2d) {
        DevBuf keys_sorted;
        auto counts = bin_windows(p, b.geom, b, &keys_sorted, s);
        b.table.reserve(std::max<size_t>(b.n, 1) * gt::kTabNW * sizeof(float));
        b.heads.reserve(std::max<size_t>(b.n, 1) * 2 * sizeof(uint32_t));
        if (b.n)
            dispatch_fast<LaunchTablesNC2C>(p->KW, p->traj.as<float>(), b.list.as<uint32_t>(),
                                            keys_sorted.as<uint32_t>(), (long long)b.n, b.geom,
                                            b.table.as<float>(), b.heads.as<uint32_t>(), mfma_weight_scale(p), s);
        GT_HIP(hipGetLastError());
        make_items(p, b, counts, segment_length(p, counts, b.n, true), true, s);
        keys_sorted.release();
    } else if (p->fast3d) {
        DevBuf keys_sorted;
        auto counts = bin_windows(p, b.geom, b, &keys_sorted, s);
        b.table.reserve(std::max<size_t>(b.n, 1) * gt::k3NW * sizeof(float));
        b.heads.reserve(std::max<size_t>(b.n, 1) * 2 * sizeof(uint32_t));
        if (b.n)
            LaunchTablesNC2C3::run(p->traj.as<float>(), b.list.as<uint32_t>(), keys_sorted.as<uint32_t>(),
                                   (long long)b.n, b.geom, b.table.as<float>(), b.heads.as<uint32_t>(), s);
        GT_HIP(hipGetLastError());
        make_items(p, b, counts, segment_length(p, counts, b.n, true), true, s);
        keys_sorted.release();
    } else {
        b.geom = p->geom;
        // double plans bin float copies of the positions with a widened window,
        // so every tile the double-precision window touches is listed
        gt::GridGeom gbin = p->geom;
        if (p->dbl) gbin.radius += gt::kBinMargin;
        auto counts = bin_windows(p, gbin, b, nullptr, s);
        make_items(p, b, counts, segment_length(p, counts, b.n, true), true, s);
    }
}

void build_c2nc(gt_nufft_plan* p, hipStream_t s) {
    if (!p->fast2d && !p->fast3d) return;  // the generic C2NC gathers per sample without binning
    Binning& b = p->bc;
    auto counts = bin_home(p, b.geom, b, s);
    if (p->fast2d) {
        b.table.reserve(std::max<size_t>(b.n, 1) * gt::kTabDw * 4);
        dispatch_fast<LaunchTablesC2NC>(p->KW, p->traj.as<float>(), b.list.as<uint32_t>(), (long long)b.n,
                                        b.geom, b.table.as<uint32_t>(), s);
        b.wtab.release();
        Binning& m = p->bm;
        if (gridding_on_matrix_cores()) {
            // the matrix-core C2NC (17 coils or more) has its own binning, its
            // home tiles wider, and per-sample weight rows
            auto cm = bin_home(p, m.geom, m, s);
            m.wtab.reserve(std::max<size_t>(m.n, 1) * 32 * sizeof(float));
            const float sx = kb_axis_scale(p, 0), sy = kb_axis_scale(p, 1);
            if (m.n)
                dispatch_fast<LaunchWTableC2NC>(p->KW, p->traj.as<float>(), m.list.as<uint32_t>(), (long long)m.n,
                                                m.geom, sx, sy, m.wtab.as<float>(), s);
            GT_HIP(hipGetLastError());
            make_items(p, m, cm, segment_length(p, cm, m.n, false), false, s);
        } else {
            m.release();
        }
    } else if (p->c2nc3_mfma) {
        b.table.release();
        b.wtab.reserve(std::max<size_t>(b.n, 1) * size_t(c2nc3_table_floats(p->c2nc3_ey)) * sizeof(float));
        const float sx = kb_axis_scale(p, 0), sy = kb_axis_scale(p, 1);
        if (b.n)
            dispatch_fast<LaunchWTableC2NC3>(p->KW, p->traj.as<float>(), b.list.as<uint32_t>(), (long long)b.n,
                                             b.geom, sx, sy, b.wtab.as<float>(), p->c2nc3_ey, s);
    } else {
        b.wtab.release();
        b.table.reserve(std::max<size_t>(b.n, 1) * gt::k3CDw * 4);
        dispatch_fast<LaunchTablesC2NC3>(p->KW, p->traj.as<float>(), b.list.as<uint32_t>(), (long long)b.n,
                                         b.geom, b.table.as<uint32_t>(), s);
    }
    GT_HIP(hipGetLastError());
    if (p->c2nc3_mfma)
        make_column_items(b, counts, segment_length(p, counts, b.n, false), s);
    else
        make_items(p, b, counts, segment_length(p, counts, b.n, false), false, s);
}

// ---- compute building blocks ----------------------------------------------
struct Shape {
    bool oversampled = false;
    size_t nb = 0;  // batches including frames (F * C)
    size_t C = 0;   // batches per frame
};

Shape image_shape(const gt_nufft_plan* p, const size_t* dims, int ndim, bool allow_plain) {
    GT_REQUIRE(dims != nullptr && ndim >= p->D, GT_ERR_SHAPE, "image has fewer dimensions than the plan");
    bool os = true, plain = true;
    for (int d = 0; d < p->D; d++) {
        os &= dims[d] == p->Nos[d];
        plain &= dims[d] == p->N[d];
    }
    GT_REQUIRE(os || (plain && allow_plain), GT_ERR_SHAPE,
               allow_plain ? "image dimensions match neither the matrix size nor the oversampled size"
                           : "image dimensions do not match the oversampled matrix size");
    Shape sh;
    sh.oversampled = os;
    sh.nb = 1;
    for (int d = p->D; d < ndim; d++) sh.nb *= dims[d];
    return sh;
}

size_t elements(const size_t* dims, int ndim) {
    size_t n = 1;
    for (int i = 0; i < ndim; i++) n *= dims[i];
    return n;
}

void check_batches(const gt_nufft_plan* p, Shape& sh, size_t sample_elems) {
    // cuGriddingConvolution.cu:229-267
    GT_REQUIRE(p->M > 0 && p->F > 0, GT_ERR_NOT_PREPROCESSED, "plan has not been preprocessed");
    GT_REQUIRE(sample_elems > 0 && sample_elems % (p->M * p->F) == 0, GT_ERR_SHAPE,
               "the number of samples is not a multiple of #samples/frame x #frames");
    GT_REQUIRE(sh.nb % p->F == 0, GT_ERR_SHAPE, "image batches are not a multiple of the frame count");
    GT_REQUIRE(sample_elems / (p->M * p->F) == sh.nb / p->F, GT_ERR_SHAPE,
               "number of batches mismatch between samples and image arrays");
    sh.C = sh.nb / p->F;
}

void conv_c2nc(gt_nufft_plan* p, const void* grid, const gt::Layout& L, void* out, size_t C, int sign,
               const float* dcw, size_t dcw_len, bool accumulate, hipStream_t s) {
    GT_REQUIRE(p->prep_c2nc, GT_ERR_NOT_PREPROCESSED, "plan not preprocessed for C2NC");
    StageTimer timer(p, GT_STAGE_C2NC, s);
    const long long dl = (long long)std::max<size_t>(dcw_len, 1);
    if (p->fast2d) {
        dispatch_fast<LaunchFastC2NC>(p->KW, (const float*)grid, (float*)out, p, int(C), L, sign, dcw, dl,
                                      accumulate ? 1 : 0, s);
    } else if (p->fast3d) {
        dispatch_fast<LaunchFast3C2NC>(p->KW, (const float*)grid, (float*)out, p, int(C), L, sign, dcw, dl,
                                       accumulate ? 1 : 0, s);
    } else {
        dispatch4<LaunchC2NC>(p->D, p->KW, (const float2*)grid, (float2*)out, p->traj.as<float>(), int(p->M),
                             int(p->F), int(C), p->geom, sign, dcw, dl, accumulate ? 1 : 0, s);
    }
    GT_HIP(hipGetLastError());
}

// partials_ext: where the split tiles' partial sums go instead of the plan's
// workspace (the caller has checked its size); null: the plan's, grown here.
void conv_nc2c(gt_nufft_plan* p, const void* samples, void* grid, const gt::Layout& L, size_t C, int sign,
               const float* dcw, size_t dcw_len, bool accumulate, hipStream_t s, void* partials_ext = nullptr) {
    GT_REQUIRE(p->prep_nc2c, GT_ERR_NOT_PREPROCESSED, "plan not preprocessed for NC2C");
    Binning& b = p->bn;
    if (p->conv_type == GT_CONV_SPARSE_MATRIX) {
        // the row-normalised convolution matrix's transpose is the standard
        // gridding of samples scaled by 1 / (window weight sum)
        const long long FM = (long long)(p->M * p->F);
        if (!dcw) {
            dcw = p->row_inv.as<float>();
            dcw_len = size_t(FM);
        } else {
            const long long n = dcw_len == size_t(FM) ? FM : FM * (long long)C;
            p->sparse_dcw.reserve(size_t(n) * sizeof(float));
            hipLaunchKernelGGL(gt::sparse_dcw_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, dcw,
                               (long long)dcw_len, p->row_inv.as<float>(), p->sparse_dcw.as<float>(), FM, n);
            GT_HIP(hipGetLastError());
            dcw = p->sparse_dcw.as<float>();
            dcw_len = size_t(n);
        }
    }
    if (p->conv_type != GT_CONV_ATOMIC && b.n_partials > 0 && !partials_ext)
        p->partials.reserve(size_t(b.n_partials) * C * b.geom.tile_cells * sizeof(float2));
    float* part = partials_ext ? static_cast<float*>(partials_ext) : p->partials.as<float>();
    StageTimer timer(p, GT_STAGE_NC2C, s);
    const long long dl = (long long)std::max<size_t>(dcw_len, 1);
    if (p->conv_type == GT_CONV_ATOMIC) {
        if (!accumulate) {
            // a fill kernel, not a memset: SENSE / CG solves over this plan are
            // recorded into hipGraphs
            const long long nz = (long long)(p->G() * p->F * C);
            hipLaunchKernelGGL(gt::zero_kernel, dim3(grid_blocks(nz)), dim3(256), 0, s, (float2*)grid, nz);
        }
        dispatch4<LaunchNC2CAtomic>(p->D, p->KW, (const float2*)samples, (float2*)grid, p->traj.as<float>(),
                                   int(p->M), int(p->F), int(C), p->geom, sign, dcw, dl, s);
        GT_HIP(hipGetLastError());
        return;
    }
    if (p->fast2d) {
        dispatch_fast<LaunchFastNC2C>(p->KW, (const float*)samples, (float*)grid, part, p,
                                      int(C), L, sign, dcw, dl, accumulate ? 1 : 0, s);
        GT_HIP(hipGetLastError());
        if (b.n_red > 0) {
            const unsigned parts = (unsigned)((C * gt::kRT * gt::kRT * 2 + 255) / 256);
            hipLaunchKernelGGL(gt::nc2c_reduce2d_kernel, dim3((unsigned)b.n_red, parts), dim3(256), 0, s,
                               part, (float*)grid, b.red.as<gt::RedItem>(), int(p->F), int(C),
                               b.geom, L, sign, accumulate ? 1 : 0);
            GT_HIP(hipGetLastError());
        }
        return;
    }
    if (p->fast3d) {
        launch_fast3_nc2c((const float*)samples, (float*)grid, part, p, int(C), L, sign, dcw,
                          dl, accumulate ? 1 : 0, s);
        GT_HIP(hipGetLastError());
        if (b.n_red > 0) {
            const unsigned parts = (unsigned)((C * b.geom.tile_cells * 2 + 255) / 256);
            hipLaunchKernelGGL(gt::nc2c_reduce3d_kernel, dim3((unsigned)b.n_red, parts), dim3(256), 0, s,
                               part, (float*)grid, b.red.as<gt::RedItem>(), int(p->F), int(C),
                               b.geom, L, sign, accumulate ? 1 : 0);
            GT_HIP(hipGetLastError());
        }
        return;
    }
    const int CC = coil_chunk(p);
    dispatch4<LaunchNC2CTile>(p->D, p->KW, (const float2*)samples, (float2*)grid, reinterpret_cast<float2*>(part),
                             p->traj.as<float>(), b.list.as<uint32_t>(), b.items.as<gt::WorkItem>(), b.n_items,
                             int(p->M), int(p->F), int(C), CC, b.geom, sign, dcw, dl, accumulate ? 1 : 0, s);
    GT_HIP(hipGetLastError());
    if (b.n_red > 0) {
        hipLaunchKernelGGL(gt::nc2c_reduce_kernel, dim3((unsigned)b.n_red), dim3(256), 0, s,
                           reinterpret_cast<float2*>(part), (float2*)grid, b.red.as<gt::RedItem>(), int(p->F), int(C),
                           b.geom, sign, accumulate ? 1 : 0);
        GT_HIP(hipGetLastError());
    }
}

// Whole-grid pad (or copy) + deapodize + centring sign + scale, caller layout.
void pad_deapod(gt_nufft_plan* p, const void* img, void* grid, size_t nb, bool in_os, float scale,
                hipStream_t s) {
    gt::PadCrop pc = padcrop(p, true, in_os, scale, 1);
    const long long rows = (long long)(nb * p->G() / p->Nos[0]);
    StageTimer timer(p, GT_STAGE_PAD, s);
    hipLaunchKernelGGL(gt::pad_deapod_kernel, dim3((unsigned)((rows + gt::kRowsPerBlock - 1) / gt::kRowsPerBlock)),
                       dim3(64, gt::kRowsPerBlock), 0, s, (const float2*)img, (float2*)grid, (long long)nb, p->geom,
                       pc);
    GT_HIP(hipGetLastError());
}

void crop_deapod(gt_nufft_plan* p, const void* grid, void* img, size_t nb, bool out_os, float scale,
                 hipStream_t s) {
    gt::PadCrop pc = padcrop(p, false, out_os, scale, 1);
    const long long rows = (long long)(nb * (out_os ? p->G() / p->Nos[0] : p->Nimg() / p->N[0]));
    StageTimer timer(p, GT_STAGE_CROP, s);
    hipLaunchKernelGGL(gt::crop_deapod_kernel, dim3((unsigned)((rows + gt::kRowsPerBlock - 1) / gt::kRowsPerBlock)),
                       dim3(64, gt::kRowsPerBlock), 0, s, (const float2*)grid, (float2*)img, (long long)nb, p->geom,
                       pc);
    GT_HIP(hipGetLastError());
}

// ---- double-precision plans: the generic path in double ---------------------
gt::PadCropD padcrop_d(const gt_nufft_plan* p, bool pad_side, bool in_os, double scale, int sign) {
    gt::PadCropD pc{};
    for (int d = 0; d < 4; d++) {
        pc.N[d] = d < p->D ? int(p->N[d]) : 1;
        if (d < p->D) {
            pc.off[d] = pad_side ? int(p->Nos[d] / 2 - p->N[d] / 2) : int((p->Nos[d] - p->N[d]) >> 1);
            pc.filt[d] = p->filt_d[d].as<double>();
        } else {
            pc.off[d] = 0;
            pc.filt[d] = nullptr;
        }
    }
    pc.scale = scale;
    pc.sign = sign;
    pc.in_os = in_os ? 1 : 0;
    return pc;
}

void pad_deapod_d(gt_nufft_plan* p, const void* img, void* grid, size_t nb, bool in_os, double scale,
                  hipStream_t s) {
    const gt::PadCropD pc = padcrop_d(p, true, in_os, scale, 1);
    const long long rows = (long long)(nb * p->G() / p->Nos[0]);
    StageTimer timer(p, GT_STAGE_PAD, s);
    hipLaunchKernelGGL(gt::pad_deapod_d_kernel, dim3((unsigned)((rows + gt::kRowsPerBlock - 1) / gt::kRowsPerBlock)),
                       dim3(64, gt::kRowsPerBlock), 0, s, (const double2*)img, (double2*)grid, (long long)nb, p->geom,
                       pc);
    GT_HIP(hipGetLastError());
}

void crop_deapod_d(gt_nufft_plan* p, const void* grid, void* img, size_t nb, bool out_os, double scale, int sign,
                   hipStream_t s) {
    const gt::PadCropD pc = padcrop_d(p, false, out_os, scale, sign);
    const long long rows = (long long)(nb * (out_os ? p->G() / p->Nos[0] : p->Nimg() / p->N[0]));
    StageTimer timer(p, GT_STAGE_CROP, s);
    hipLaunchKernelGGL(gt::crop_deapod_d_kernel, dim3((unsigned)((rows + gt::kRowsPerBlock - 1) / gt::kRowsPerBlock)),
                       dim3(64, gt::kRowsPerBlock), 0, s, (const double2*)grid, (double2*)img, (long long)nb, p->geom,
                       pc);
    GT_HIP(hipGetLastError());
}

void conv_c2nc_d(gt_nufft_plan* p, const void* grid, void* out, size_t C, int sign, const double* dcw,
                 size_t dcw_len, bool accumulate, hipStream_t s) {
    GT_REQUIRE(p->prep_c2nc, GT_ERR_NOT_PREPROCESSED, "plan not preprocessed for C2NC");
    StageTimer timer(p, GT_STAGE_C2NC, s);
    dispatch4<LaunchC2NCD>(p->D, p->KW, (const double2*)grid, (double2*)out, p->traj_d.as<double>(), int(p->M),
                          int(p->F), int(C), p->geom, sign, dcw, (long long)std::max<size_t>(dcw_len, 1),
                          accumulate ? 1 : 0, s);
    GT_HIP(hipGetLastError());
}

void conv_nc2c_d(gt_nufft_plan* p, const void* samples, void* grid, size_t C, int sign, const double* dcw,
                 size_t dcw_len, bool accumulate, hipStream_t s, void* partials_ext = nullptr) {
    GT_REQUIRE(p->prep_nc2c, GT_ERR_NOT_PREPROCESSED, "plan not preprocessed for NC2C");
    Binning& b = p->bn;
    const int TC = std::max(1, b.geom.tile_cells);
    if (b.n_partials > 0 && !partials_ext) p->partials.reserve(size_t(b.n_partials) * C * TC * sizeof(double2));
    double2* part = partials_ext ? static_cast<double2*>(partials_ext) : p->partials.as<double2>();
    StageTimer timer(p, GT_STAGE_NC2C, s);
    // LDS budget per one-wave workgroup: 16 KiB of re/im accumulators
    const int CC = std::max(1, std::min(16384 / (16 * TC), 32));
    const long long dl = (long long)std::max<size_t>(dcw_len, 1);
    dispatch4<LaunchNC2CTileD>(p->D, p->KW, (const double2*)samples, (double2*)grid, part,
                              p->traj_d.as<double>(), b.list.as<uint32_t>(), b.items.as<gt::WorkItem>(), b.n_items,
                              int(p->M), int(p->F), int(C), CC, b.geom, sign, dcw, dl, accumulate ? 1 : 0, s);
    GT_HIP(hipGetLastError());
    if (b.n_red > 0) {
        hipLaunchKernelGGL(gt::nc2c_reduce_d_kernel, dim3((unsigned)b.n_red), dim3(256), 0, s,
                           part, (double2*)grid, b.red.as<gt::RedItem>(), int(p->F), int(C),
                           b.geom, sign, accumulate ? 1 : 0);
        GT_HIP(hipGetLastError());
    }
}

// compute() of a double-precision plan (NFFT.hpp:104-238, modes as in compute_impl);
// dcw points to doubles.
void compute_impl_d(gt_nufft_plan* p, const void* in, void* out, const Shape& sh, const double* dcw, size_t dcw_len,
                    int mode, hipStream_t s) {
    const size_t G = p->G();
    const double scale = double(centring_global_sign(p)) / std::sqrt(double(G));
    switch (mode) {
    case GT_FORWARDS_C2NC:
    case GT_BACKWARDS_C2NC:
        p->grid.reserve(sh.nb * G * sizeof(double2));
        pad_deapod_d(p, in, p->grid.ptr, sh.nb, sh.oversampled, scale, s);
        run_fft(p, p->grid.ptr, sh.nb, mode == GT_FORWARDS_C2NC, s);
        conv_c2nc_d(p, p->grid.ptr, out, sh.C, 1, dcw, dcw_len, false, s);
        break;
    case GT_BACKWARDS_NC2C:
    case GT_FORWARDS_NC2C: {
        void* g = out;
        if (!sh.oversampled) {
            p->grid.reserve(sh.nb * G * sizeof(double2));
            g = p->grid.ptr;
        }
        conv_nc2c_d(p, in, g, sh.C, 1, dcw, dcw_len, false, s);
        run_fft(p, g, sh.nb, mode == GT_FORWARDS_NC2C, s);
        crop_deapod_d(p, g, out, sh.nb, sh.oversampled, scale, 1, s);
        break;
    }
    }
}

gt::RowGeom row_geom(const gt_nufft_plan* p, size_t nb, bool pad_side, float scale) {
    gt::RowGeom rg{};
    rg.Nx = int(p->N[0]);
    rg.Ny = int(p->N[1]);
    rg.Ox = int(p->Nos[0]);
    rg.Oy = int(p->Nos[1]);
    rg.offx = pad_side ? int(p->Nos[0] / 2 - p->N[0] / 2) : int((p->Nos[0] - p->N[0]) >> 1);
    rg.offy = pad_side ? int(p->Nos[1] / 2 - p->N[1] / 2) : int((p->Nos[1] - p->N[1]) >> 1);
    rg.nb = int(nb);
    rg.fx = p->filt[0].as<float>();
    rg.fy = p->filt[1].as<float>();
    rg.scale = scale;
    rg.sign = 1;
    return rg;
}

// ---- fused band FFT launchers (gt_fft.h) -------------------------------------
template <int L>
struct LaunchRowsPad {
    static void run(const float2* img, float2* grid, const gt::FftGeom& fg, const float2* tw, float dir,
                    hipStream_t s) {
        const unsigned blocks = unsigned((size_t(fg.nb) * fg.Ny + gt::kFftRows - 1) / gt::kFftRows);
        hipLaunchKernelGGL(gt::fft_rows_pad_kernel<L>, dim3(blocks), dim3(gt::kFftRowThreads * gt::kFftRows), 0, s,
                           img, grid, fg, tw, dir);
    }
};
template <int L>
struct LaunchRowsCrop {
    static void run(const float2* grid, float2* img, const gt::FftGeom& fg, const float2* tw, float dir,
                    hipStream_t s) {
        const unsigned blocks = unsigned((size_t(fg.nb) * fg.Ny + gt::kFftRows - 1) / gt::kFftRows);
        hipLaunchKernelGGL(gt::fft_rows_crop_kernel<L>, dim3(blocks), dim3(gt::kFftRowThreads * gt::kFftRows), 0,
                           s, grid, img, fg, tw, dir);
    }
};
template <int L>
struct LaunchColsBandIn {
    static void run(float2* grid, int Ox, const gt::FftGeom& fg, const float2* tw, float dir, hipStream_t s) {
        constexpr int W = gt::fft_col_width<L>();
        size_t blocks = size_t(fg.nb) * size_t((Ox + W - 1) / W);
        if (W == 8) blocks = (blocks + 15) / 16 * 16;  // whole groups of the XCD-paired order
        hipLaunchKernelGGL((gt::fft_cols_kernel<L, true, false>), dim3(unsigned(blocks)), dim3(256), 0, s, grid, Ox, fg, tw,
                           dir);
    }
};
template <int L>
struct LaunchColsBandOut {
    static void run(float2* grid, int Ox, const gt::FftGeom& fg, const float2* tw, float dir, hipStream_t s) {
        constexpr int W = gt::fft_col_width<L>();
        size_t blocks = size_t(fg.nb) * size_t((Ox + W - 1) / W);
        if (W == 8) blocks = (blocks + 15) / 16 * 16;  // whole groups of the XCD-paired order
        hipLaunchKernelGGL((gt::fft_cols_kernel<L, false, true>), dim3(unsigned(blocks)), dim3(256), 0, s, grid, Ox, fg, tw,
                           dir);
    }
};

template <int L>
struct LaunchRows3Pad {
    static void run(const float2* img, float2* grid, const gt::Fft3Geom& fg, const float2* tw, float dir,
                    hipStream_t s) {
        const size_t rows = size_t(fg.nb) * fg.Nz * fg.Ny;
        hipLaunchKernelGGL(gt::fft3_rows_pad_kernel<L>, dim3(unsigned((rows + gt::kFftRows - 1) / gt::kFftRows)),
                           dim3(gt::kFftRowThreads * gt::kFftRows), 0, s, img, grid, fg, tw, dir);
    }
};
template <int L>
struct LaunchRows3Crop {
    static void run(const float2* grid, float2* img, const gt::Fft3Geom& fg, const float2* tw, float dir,
                    hipStream_t s) {
        const size_t rows = size_t(fg.nb) * fg.Nz * fg.Ny;
        hipLaunchKernelGGL(gt::fft3_rows_crop_kernel<L>, dim3(unsigned((rows + gt::kFftRows - 1) / gt::kFftRows)),
                           dim3(gt::kFftRowThreads * gt::kFftRows), 0, s, grid, img, fg, tw, dir);
    }
};
size_t lines_blocks(const gt::FftLines& fl, int W) {
    size_t blocks = size_t(fl.nbat) * size_t((fl.ncols + W - 1) / W);
    if (W == 8) blocks = (blocks + 15) / 16 * 16;  // whole groups of the XCD-paired order
    return blocks;
}
// Lines per workgroup of the 3D y and z passes: 32 (256-byte runs of lines
// that lie N_os[0] or a whole plane apart) where the block's LDS stays below
// 64 KB, else 16.  128^3 -> 192^3, 8 coils: FFT stage 0.86 ms with 8 lines
// per workgroup, 0.73 ms with 16, 0.70 ms with 32.
template <int L>
constexpr int lines_width() { return L <= 240 ? 32 : 16; }
template <int L, bool BI, bool BO>
void launch_lines(float2* grid, const gt::FftLines& fl, const float2* tw, float dir, hipStream_t s) {
    constexpr int W = lines_width<L>();
    hipLaunchKernelGGL((gt::fft_lines_kernel<L, BI, BO, W>), dim3(unsigned(lines_blocks(fl, W))), dim3(256), 0, s, grid,
                       fl, tw, dir);
}
template <int L>
struct LaunchLinesBandIn {
    static void run(float2* grid, const gt::FftLines& fl, const float2* tw, float dir, hipStream_t s) {
        launch_lines<L, true, false>(grid, fl, tw, dir, s);
    }
};
template <int L>
struct LaunchLinesBandOut {
    static void run(float2* grid, const gt::FftLines& fl, const float2* tw, float dir, hipStream_t s) {
        launch_lines<L, false, true>(grid, fl, tw, dir, s);
    }
};

gt::FftGeom fft_geom(const gt_nufft_plan* p, size_t nb, bool pad_side, float scale) {
    const gt::RowGeom rg = row_geom(p, nb, pad_side, scale);
    gt::FftGeom fg{};
    fg.Nx = rg.Nx;
    fg.Ny = rg.Ny;
    fg.offx = rg.offx;
    fg.offy = rg.offy;
    fg.nb = rg.nb;
    fg.fx = rg.fx;
    fg.fy = rg.fy;
    fg.scale = rg.scale;
    return fg;
}

gt::Fft3Geom fft3_geom(const gt_nufft_plan* p, size_t nb, const gt::Layout& L, bool pad_side, float scale) {
    const gt::PadCrop pc = padcrop(p, pad_side, false, scale, 1);
    gt::Fft3Geom fg{};
    fg.sb = L.sb;
    fg.sy = L.sy;
    fg.sz = L.sz;
    fg.Nx = pc.N[0];
    fg.Ny = pc.N[1];
    fg.Nz = pc.N[2];
    fg.Ox = int(p->Nos[0]);
    fg.Oy = int(p->Nos[1]);
    fg.Oz = int(p->Nos[2]);
    fg.offx = pc.off[0];
    fg.offy = pc.off[1];
    fg.offz = pc.off[2];
    fg.nb = int(nb);
    fg.fx = pc.filt[0];
    fg.fy = pc.filt[1];
    fg.fz = pc.filt[2];
    fg.scale = scale;
    return fg;
}

// 3D plans with the band transforms of gt_fft.h around the gridding (image of
// the matrix size): forwards x (with pad and deapodization), y over the planes
// of the z band, z over every line; backwards the same in reverse, with the
// crop in the x pass.  The grid between them is [z][y][batch][x], so the coil
// rows that the gridding kernels stage together for one (y, z) lie within
// batches * N_os[0] cells of each other instead of a whole grid apart.
void compute_fused3(gt_nufft_plan* p, const void* in, void* out, const Shape& sh, const float* dcw,
                    size_t dcw_len, int mode, float scale, hipStream_t s) {
    const size_t nb = sh.nb;
    GT_REQUIRE(nb * p->G() < (size_t(1) << 40) && nb * p->Nimg() / p->N[0] < (size_t(1) << 31) &&
                   nb * p->N[2] * size_t((p->Nos[0] + 7) / 8) < (size_t(1) << 31) &&
                   nb * p->Nos[0] * p->Nos[1] < (size_t(1) << 31),
               GT_ERR_UNSUPPORTED, "too many batches for one launch");
    p->grid.reserve(nb * p->G() * sizeof(float2));
    float2* g = p->grid.as<float2>();
    const gt::Layout L = internal_layout(p, nb);  // [z][y][batch][x]
    const int Ox = int(p->Nos[0]);
    const float2* tw[3] = {p->twiddle[0].as<float2>(), p->twiddle[1].as<float2>(), p->twiddle[2].as<float2>()};
    const bool c2nc = mode == GT_FORWARDS_C2NC || mode == GT_BACKWARDS_C2NC;
    const gt::Fft3Geom fg = fft3_geom(p, nb, L, c2nc, scale);
    const int Oy = int(p->Nos[1]), Oz = int(p->Nos[2]);
    // y lines: batch bb = iz * nb + b of the z band, Ox lines each; z lines:
    // one batch of every (y, batch, x) line
    const gt::FftLines ly{int(nb) * fg.Nz, int(nb), fg.offz, 0, Ox, L.sz, L.sb, L.sy, fg.offy, fg.offy + fg.Ny};
    const gt::FftLines lz{1, 1, 0, 0, int(L.sz), 0, 0, L.sz, fg.offz, fg.offz + fg.Nz};
    if (c2nc) {
        const float dir = mode == GT_FORWARDS_C2NC ? -1.0f : 1.0f;
        {
            StageTimer timer(p, GT_STAGE_FFT, s);
            dispatch_fft_len<LaunchRows3Pad>(Ox, (const float2*)in, g, fg, tw[0], dir, s);
            GT_HIP(hipGetLastError());
            dispatch_fft_len<LaunchLinesBandIn>(Oy, g, ly, tw[1], dir, s);
            GT_HIP(hipGetLastError());
            dispatch_fft_len<LaunchLinesBandIn>(Oz, g, lz, tw[2], dir, s);
            GT_HIP(hipGetLastError());
        }
        conv_c2nc(p, g, L, out, sh.C, 1, dcw, dcw_len, false, s);
    } else {
        const float dir = mode == GT_FORWARDS_NC2C ? -1.0f : 1.0f;
        conv_nc2c(p, in, g, L, sh.C, 1, dcw, dcw_len, false, s);
        {
            StageTimer timer(p, GT_STAGE_FFT, s);
            dispatch_fft_len<LaunchLinesBandOut>(Oz, g, lz, tw[2], dir, s);
            GT_HIP(hipGetLastError());
            dispatch_fft_len<LaunchLinesBandOut>(Oy, g, ly, tw[1], dir, s);
            GT_HIP(hipGetLastError());
            dispatch_fft_len<LaunchRows3Crop>(Ox, (const float2*)g, (float2*)out, fg, tw[0], dir, s);
            GT_HIP(hipGetLastError());
        }
    }
}

// 2D hot path with the fused band transforms: pad + deapodize + row FFT over
// the image rows, then column FFTs (C2NC side); column FFTs writing only the
// cropped rows, then row FFT + crop + deapodize (NC2C side).
// csm (FORWARDS_C2NC only): coil maps [N..., C]; `in` is then one image per
// frame, multiplied by each coil's map inside the row pass.
// grid_ext / partials_ext: caller-provided grid (nb * G complex) and split-tile
// partial sums; null: the plan's workspaces, grown here.
void compute_fast2d_fused(gt_nufft_plan* p, const void* in, void* out, const Shape& sh, const float* dcw,
                          size_t dcw_len, int mode, float scale, hipStream_t s, const float2* csm = nullptr,
                          float2* grid_ext = nullptr, void* partials_ext = nullptr) {
    const size_t nb = sh.nb;
    const int Ox = int(p->Nos[0]), Oy = int(p->Nos[1]);
    GT_REQUIRE(nb * p->N[1] < (size_t(1) << 31) && nb * size_t((Ox + 15) / 16) < (size_t(1) << 31),
               GT_ERR_UNSUPPORTED, "too many batches for one launch");
    if (!grid_ext) p->grid.reserve(nb * p->G() * sizeof(float2));
    float2* g = grid_ext ? grid_ext : p->grid.as<float2>();
    const gt::Layout L = internal_layout(p, nb);
    const float2* twx = p->twiddle[0].as<float2>();
    const float2* twy = p->twiddle[1].as<float2>();
    if (mode == GT_FORWARDS_C2NC || mode == GT_BACKWARDS_C2NC) {
        const float dir = mode == GT_FORWARDS_C2NC ? -1.0f : 1.0f;
        gt::FftGeom fg = fft_geom(p, nb, true, scale);
        fg.csm = csm;
        fg.F = int(p->F);
        {
            StageTimer timer(p, GT_STAGE_FFT, s);
            dispatch_fft_len<LaunchRowsPad>(Ox, (const float2*)in, g, fg, twx, dir, s);
            GT_HIP(hipGetLastError());
            dispatch_fft_len<LaunchColsBandIn>(Oy, g, Ox, fg, twy, dir, s);
            GT_HIP(hipGetLastError());
        }
        conv_c2nc(p, g, L, out, sh.C, 1, dcw, dcw_len, false, s);
    } else {
        const float dir = mode == GT_FORWARDS_NC2C ? -1.0f : 1.0f;
        const gt::FftGeom fg = fft_geom(p, nb, false, scale);
        conv_nc2c(p, in, g, L, sh.C, 1, dcw, dcw_len, false, s, partials_ext);
        {
            StageTimer timer(p, GT_STAGE_FFT, s);
            dispatch_fft_len<LaunchColsBandOut>(Oy, g, Ox, fg, twy, dir, s);
            GT_HIP(hipGetLastError());
            dispatch_fft_len<LaunchRowsCrop>(Ox, (const float2*)g, (float2*)out, fg, twx, dir, s);
            GT_HIP(hipGetLastError());
        }
    }
}

// 2D hot path, image side of size N: internal layout [y][batch][x].
void compute_fast2d(gt_nufft_plan* p, const void* in, void* out, const Shape& sh, const float* dcw, size_t dcw_len,
                    int mode, float scale, hipStream_t s) {
    if (p->fft_fused && fft2d_strategy() == 2) {
        compute_fast2d_fused(p, in, out, sh, dcw, dcw_len, mode, scale, s);
        return;
    }
    const size_t nb = sh.nb;
    const size_t Ox = p->Nos[0];
    p->grid.reserve(nb * p->G() * sizeof(float2));
    float2* g = p->grid.as<float2>();
    const gt::Layout L = internal_layout(p, nb);
    if (mode == GT_FORWARDS_C2NC || mode == GT_BACKWARDS_C2NC) {
        const bool fwd = mode == GT_FORWARDS_C2NC;
        gt::RowGeom rg = row_geom(p, nb, true, scale);
        {
            StageTimer timer(p, GT_STAGE_PAD, s);
            hipLaunchKernelGGL(gt::pad_rows2d_kernel, dim3((unsigned)(p->Nos[1] * nb)), dim3(256), 0, s,
                               (const float2*)in, g, rg);
            GT_HIP(hipGetLastError());
        }
        {
            StageTimer timer(p, GT_STAGE_FFT, s);
            if (fft2d_strategy() == 1) {
                exec_fft(p, FFT_GRID2D, nb, g, fwd, s);
            } else {
                // rows outside the padded image are zero: transform only the image rows
                exec_fft(p, FFT_ROWS, p->N[1] * nb, g + size_t(rg.offy) * nb * Ox, fwd, s);
                exec_fft(p, FFT_COLS, nb, g, fwd, s);
            }
        }
        conv_c2nc(p, g, L, out, sh.C, 1, dcw, dcw_len, false, s);
    } else {
        const bool fwd = mode == GT_FORWARDS_NC2C;
        gt::RowGeom rg = row_geom(p, nb, false, scale);
        conv_nc2c(p, in, g, L, sh.C, 1, dcw, dcw_len, false, s);
        {
            StageTimer timer(p, GT_STAGE_FFT, s);
            if (fft2d_strategy() == 1) {
                exec_fft(p, FFT_GRID2D, nb, g, fwd, s);
            } else {
                exec_fft(p, FFT_COLS, nb, g, fwd, s);
                // only the rows kept by the crop are needed
                exec_fft(p, FFT_ROWS, p->N[1] * nb, g + size_t(rg.offy) * nb * Ox, fwd, s);
            }
        }
        {
            StageTimer timer(p, GT_STAGE_CROP, s);
            hipLaunchKernelGGL(gt::crop_rows2d_kernel, dim3((unsigned)(p->N[1] * nb)), dim3(256), 0, s,
                               (const float2*)g, (float2*)out, rg);
            GT_HIP(hipGetLastError());
        }
    }
}

void compute_impl(gt_nufft_plan* p, const void* in, const size_t* in_dims, int in_ndim, void* out,
                  const size_t* out_dims, int out_ndim, const float* dcw, size_t dcw_len, int mode,
                  hipStream_t s) {
    GT_REQUIRE(in && out, GT_ERR_INVALID_ARGUMENT, "null input or output");
    GT_REQUIRE(mode >= 0 && mode <= 3, GT_ERR_INVALID_ARGUMENT, "unknown NFFT_comp_mode");
    GT_REQUIRE(!dcw || dcw_len > 0, GT_ERR_INVALID_ARGUMENT, "dcw given with zero length");
    const bool image_in = (mode == GT_FORWARDS_C2NC || mode == GT_BACKWARDS_C2NC);
    Shape sh = image_in ? image_shape(p, in_dims, in_ndim, true) : image_shape(p, out_dims, out_ndim, true);
    const size_t selems = image_in ? elements(out_dims, out_ndim) : elements(in_dims, in_ndim);
    check_batches(p, sh, selems);
    set_device(p);
    if (p->dbl) {
        compute_impl_d(p, in, out, sh, reinterpret_cast<const double*>(dcw), dcw_len, mode, s);
        return;
    }
    const size_t G = p->G();
    const float gs = centring_global_sign(p);
    const float scale = gs / std::sqrt(float(G));  // do_scale=true (NFFT.h:114-115, NFFT.hpp:205-238)
    if (p->fast2d && !sh.oversampled) {
        compute_fast2d(p, in, out, sh, dcw, dcw_len, mode, scale, s);
        return;
    }
    // the band transforms keep the grid as [z][y][batch][x], which only the
    // 3D hot path's gridding kernels address (the generic and atomic kernels
    // take the caller layout)
    if (p->fast3d && p->fft_fused3 && !sh.oversampled && fft2d_strategy() == 2) {
        compute_fused3(p, in, out, sh, dcw, dcw_len, mode, scale, s);
        return;
    }
    const gt::Layout L = caller_layout(p);
    switch (mode) {
    case GT_FORWARDS_C2NC:
    case GT_BACKWARDS_C2NC: {
        // FORWARDS_C2NC: deapodize(image) -> FFT forwards -> C2NC (NFFT.hpp:205-211).
        // BACKWARDS_C2NC: the adjoint of FORWARDS_NC2C (deapodize(fourier) -> FFT backwards -> C2NC).
        p->grid.reserve(sh.nb * G * sizeof(float2));
        pad_deapod(p, in, p->grid.ptr, sh.nb, sh.oversampled, scale, s);
        run_fft(p, p->grid.ptr, sh.nb, mode == GT_FORWARDS_C2NC, s);
        conv_c2nc(p, p->grid.ptr, L, out, sh.C, 1, dcw, dcw_len, false, s);
        break;
    }
    case GT_BACKWARDS_NC2C:
    case GT_FORWARDS_NC2C: {
        // BACKWARDS_NC2C: dcw -> NC2C -> FFT backwards -> deapodize -> crop (NFFT.hpp:223-229).
        // FORWARDS_NC2C: dcw -> NC2C -> FFT forwards -> deapodize(fourier) -> crop (NFFT.hpp:214-220).
        void* g = out;
        if (!sh.oversampled) {
            p->grid.reserve(sh.nb * G * sizeof(float2));
            g = p->grid.ptr;
        }
        conv_nc2c(p, in, g, L, sh.C, 1, dcw, dcw_len, false, s);
        run_fft(p, g, sh.nb, mode == GT_FORWARDS_NC2C, s);
        crop_deapod(p, g, out, sh.nb, sh.oversampled, scale, s);
        break;
    }
    }
}

}  // namespace

// ---- fused gridding of many 2D problems ------------------------------------
struct gt_host::FusedGridding {
    struct Group {
        gt_nufft_plan* rep = nullptr;  // one plan of the group: filters, twiddles, geometry
        size_t nb = 0;                 // batches of the group
        DevBuf grid;                   // [Oy][nb][Ox]
        DevBuf coil;                   // [nb][N1][N0]
        DevBuf img_b, csm_b;           // per-batch image and coil-map pointers
    };
    std::vector<FusedProblem> probs;
    std::vector<int> group_of;   // problem -> group
    std::vector<size_t> boff;    // problem -> its first batch in the group
    std::vector<Group> groups;
    DevBuf samples;              // internal samples of every problem
    std::vector<size_t> soff;    // problem -> its offset into samples (complex elements)
    DevBuf nc2c_data, nc2c_int, c2nc;  // gt::MfGridArgs per problem
    DevBuf nc2c_units, c2nc_units, red_units, reds;
    int n_nc2c = 0, n_c2nc = 0, n_red = 0, red_parts = 0;
    int dm = 0;                  // dcw mode of the NC2C launches
    int KW = 0;
};

namespace {

constexpr int kFusedCW = 16;  // coils per wave / workgroup of the fused gridding

template <int KW>
struct LaunchFusedC2NC {
    static void run(const gt::MfGridArgs* args, const gt::GridUnit* units, int n, hipStream_t s) {
        hipLaunchKernelGGL((gt::c2nc_mfma2d_batch_kernel<KW, kFusedCW>), dim3((unsigned)n), dim3(256),
                           size_t(gt::c2nc_mfma_lds_bytes(KW, kFusedCW)), s, args, units);
    }
};

// Units of one problem's matrix-core gridding in the order of its own launch
// (block b = (item / 8) 8 chunks + chunk 8 + item % 8, so b mod 8 = item mod
// 8: the chunks of an item share its XCD); a problem's run is a multiple of 8
// long, so the order survives concatenation.
void append_units(std::vector<gt::GridUnit>& u, int prob, int n_items, int nchunks) {
    const int nblk = (n_items + 7) / 8 * 8 * nchunks, grp = 8 * nchunks;
    for (int b = 0; b < nblk; b++) {
        const int item = (b / grp) * 8 + (b & 7), chunk = (b / 8) % nchunks;
        u.push_back(item < n_items ? gt::GridUnit{prob, item, chunk, 0} : gt::GridUnit{-1, 0, 0, 0});
    }
}

}  // namespace

std::string gt_host::fused_unsupported(const FusedProblem& q) {
    const gt_nufft_plan* p = q.plan;
    if (!p) return "null plan";
    if (p->dbl || p->D != 2) return "the fused batch takes float 2D plans";
    if (!(p->fast2d && p->fft_fused && fft2d_strategy() == 2))
        return "the fused batch takes plans whose oversampled sizes are fused FFT lengths (96 ... 384)";
    if (p->conv_type != GT_CONV_STANDARD) return "the fused batch takes STANDARD convolution";
    if (!gridding_on_matrix_cores()) return "the fused batch runs on the matrix cores (GT_NUFFT_MFMA=0 is set)";
    if (!(p->prep_c2nc && p->prep_nc2c) || !p->bm.wtab.ptr) return "plan not preprocessed for both directions";
    if (q.C < 1 || !q.csm || !q.image || !q.data) return "missing coil maps, image or data";
    return "";
}

gt_host::FusedGridding* gt_host::fused_create(const std::vector<FusedProblem>& probs, hipStream_t s) {
    GT_REQUIRE(!probs.empty(), GT_ERR_INVALID_ARGUMENT, "no problems");
    const bool with_dcw = probs[0].dcw != nullptr;
    for (const auto& q : probs) {
        const std::string why = fused_unsupported(q);
        GT_REQUIRE(why.empty(), GT_ERR_UNSUPPORTED, why);
        GT_REQUIRE((q.dcw != nullptr) == with_dcw, GT_ERR_UNSUPPORTED,
                   "the fused batch takes a dcw for every problem or for none");
        GT_REQUIRE(q.plan->KW == probs[0].plan->KW && q.plan->device == probs[0].plan->device, GT_ERR_UNSUPPORTED,
                   "the fused batch takes one kernel width and one device");
    }
    {
        std::vector<const gt_nufft_plan*> seen;
        for (const auto& q : probs) seen.push_back(q.plan);
        std::sort(seen.begin(), seen.end());
        GT_REQUIRE(std::adjacent_find(seen.begin(), seen.end()) == seen.end(), GT_ERR_INVALID_ARGUMENT,
                   "every problem of a fused batch needs a plan of its own (the split tiles' partial sums are "
                   "the plan's)");
    }
    auto f = std::make_unique<FusedGridding>();
    f->probs = probs;
    f->KW = probs[0].plan->KW;
    set_device(probs[0].plan);
    const int n = int(probs.size());
    std::map<std::array<size_t, 4>, int> gid;
    f->group_of.resize(size_t(n));
    f->boff.resize(size_t(n));
    for (int i = 0; i < n; i++) {
        gt_nufft_plan* p = probs[size_t(i)].plan;
        const std::array<size_t, 4> key{p->N[0], p->N[1], p->Nos[0], p->Nos[1]};
        auto it = gid.find(key);
        if (it == gid.end()) {
            it = gid.emplace(key, int(f->groups.size())).first;
            f->groups.emplace_back();
            f->groups.back().rep = p;
        }
        auto& g = f->groups[size_t(it->second)];
        f->group_of[size_t(i)] = it->second;
        f->boff[size_t(i)] = g.nb;
        g.nb += p->F * probs[size_t(i)].C;
    }
    // per-batch image / coil-map pointers of each group's forward row FFT
    for (size_t gi = 0; gi < f->groups.size(); gi++) {
        auto& g = f->groups[gi];
        const gt_nufft_plan* p = g.rep;
        GT_REQUIRE(g.nb * p->N[1] < (size_t(1) << 31) && g.nb * size_t((p->Nos[0] + 15) / 16) < (size_t(1) << 31),
                   GT_ERR_UNSUPPORTED, "too many batches for one launch");
        g.grid.reserve(g.nb * p->G() * sizeof(float2));
        g.coil.reserve(g.nb * p->Nimg() * sizeof(float2));
        std::vector<const float2*> ib(g.nb), cb(g.nb);
        for (int i = 0; i < n; i++) {
            if (f->group_of[size_t(i)] != int(gi)) continue;
            const FusedProblem& q = probs[size_t(i)];
            const size_t F = q.plan->F, nimg = q.plan->Nimg();
            for (size_t c = 0; c < q.C; c++)
                for (size_t fr = 0; fr < F; fr++) {
                    const size_t b = f->boff[size_t(i)] + c * F + fr;
                    ib[b] = static_cast<const float2*>(q.image) + fr * nimg;
                    cb[b] = static_cast<const float2*>(q.csm) + c * nimg;
                }
        }
        g.img_b.reserve(g.nb * sizeof(void*));
        g.csm_b.reserve(g.nb * sizeof(void*));
        GT_HIP(hipMemcpyAsync(g.img_b.ptr, ib.data(), g.nb * sizeof(void*), hipMemcpyHostToDevice, s));
        GT_HIP(hipMemcpyAsync(g.csm_b.ptr, cb.data(), g.nb * sizeof(void*), hipMemcpyHostToDevice, s));
        GT_HIP(hipStreamSynchronize(s));
    }
    // internal samples
    size_t total = 0;
    f->soff.resize(size_t(n));
    for (int i = 0; i < n; i++) {
        f->soff[size_t(i)] = total;
        total += probs[size_t(i)].plan->M * probs[size_t(i)].plan->F * probs[size_t(i)].C;
    }
    f->samples.reserve(std::max<size_t>(total, 1) * sizeof(float2));
    // per-problem arguments and the unit lists
    std::vector<gt::MfGridArgs> nd(static_cast<size_t>(n)), ni(static_cast<size_t>(n)), cn(static_cast<size_t>(n));
    std::vector<gt::GridUnit> un, uc, ur;
    std::vector<const gt::RedItem*> reds(size_t(n), nullptr);
    bool all_per_sample = true;
    size_t cmax = 1;
    for (int i = 0; i < n; i++) {
        const FusedProblem& q = probs[size_t(i)];
        gt_nufft_plan* p = q.plan;
        const auto& g = f->groups[size_t(f->group_of[size_t(i)])];
        const int C = int(q.C);
        const gt::Layout L = internal_layout(p, g.nb);
        float* gridp = g.grid.as<float>() + 2 * f->boff[size_t(i)] * p->Nos[0];
        float2* smp = f->samples.as<float2>() + f->soff[size_t(i)];
        if (p->bn.n_partials > 0) p->partials.reserve(size_t(p->bn.n_partials) * q.C * p->bn.geom.tile_cells * sizeof(float2));
        const long long dl = (long long)std::max<size_t>(q.dcw_len, 1);
        if (q.dcw && q.dcw_len != p->M * p->F) all_per_sample = false;
        nd[size_t(i)] = gt::MfGridArgs{static_cast<const float*>(q.data), gridp, p->partials.as<float>(),
                                       p->bn.table.as<float>(), p->bn.heads.as<uint32_t>(),
                                       p->bn.items.as<gt::WorkItem>(), int(p->M), int(p->F), C, p->bn.geom, L, 1,
                                       q.dcw, dl, 0, mfma_weight_scale(p)};
        ni[size_t(i)] = nd[size_t(i)];
        ni[size_t(i)].in = reinterpret_cast<const float*>(smp);
        cn[size_t(i)] = gt::MfGridArgs{gridp, reinterpret_cast<float*>(smp), nullptr, p->bm.wtab.as<float>(),
                                       p->bm.list.as<uint32_t>(), p->bm.items.as<gt::WorkItem>(), int(p->M),
                                       int(p->F), C, p->bm.geom, L, 1, q.dcw, dl, 0,
                                       1.0f / (kb_axis_scale(p, 0) * kb_axis_scale(p, 1))};
        const int chunks = (C + kFusedCW - 1) / kFusedCW;
        append_units(un, i, p->bn.n_items, chunks);
        append_units(uc, i, p->bm.n_items, chunks);
        reds[size_t(i)] = p->bn.red.as<gt::RedItem>();
        for (int k = 0; k < p->bn.n_red; k++) ur.push_back(gt::GridUnit{i, k, 0, 0});
        cmax = std::max(cmax, q.C);
    }
    f->dm = !with_dcw ? 0 : (all_per_sample ? 1 : 2);
    f->n_nc2c = int(un.size());
    f->n_c2nc = int(uc.size());
    f->n_red = int(ur.size());
    f->red_parts = int((cmax * gt::kRT * gt::kRT * 2 + 255) / 256);
    auto upload = [&](DevBuf& d, const void* src, size_t bytes) {
        d.reserve(std::max<size_t>(bytes, 1));
        if (bytes) GT_HIP(hipMemcpyAsync(d.ptr, src, bytes, hipMemcpyHostToDevice, s));
    };
    upload(f->nc2c_data, nd.data(), nd.size() * sizeof(gt::MfGridArgs));
    upload(f->nc2c_int, ni.data(), ni.size() * sizeof(gt::MfGridArgs));
    upload(f->c2nc, cn.data(), cn.size() * sizeof(gt::MfGridArgs));
    upload(f->nc2c_units, un.data(), un.size() * sizeof(gt::GridUnit));
    upload(f->c2nc_units, uc.data(), uc.size() * sizeof(gt::GridUnit));
    upload(f->red_units, ur.data(), ur.size() * sizeof(gt::GridUnit));
    upload(f->reds, reds.data(), reds.size() * sizeof(void*));
    GT_HIP(hipStreamSynchronize(s));
    return f.release();
}

void gt_host::fused_destroy(FusedGridding* f) { delete f; }

void gt_host::plan_retile_nc2c(gt_nufft_plan* p, int len, hipStream_t s) {
    GT_REQUIRE(p && p->fast2d && p->prep_nc2c, GT_ERR_INVALID_ARGUMENT, "plan not preprocessed on the 2D hot path");
    if (p->nc2c_len_req == len) return;
    set_device(p);
    p->nc2c_len_req = len;
    p->generation++;
    build_nc2c(p, s);
}

void gt_host::fused_forward(FusedGridding* f, hipStream_t s) {
    for (auto& g : f->groups) {
        gt_nufft_plan* p = g.rep;
        const float scale = centring_global_sign(p) / std::sqrt(float(p->G()));
        gt::FftGeom fg = fft_geom(p, g.nb, true, scale);
        fg.F = 1;
        fg.img_b = g.img_b.as<const float2*>();
        fg.csm_b = g.csm_b.as<const float2*>();
        float2* grid = g.grid.as<float2>();
        dispatch_fft_len<LaunchRowsPad>(int(p->Nos[0]), (const float2*)nullptr, grid, fg, p->twiddle[0].as<float2>(),
                                        -1.0f, s);
        GT_HIP(hipGetLastError());
        dispatch_fft_len<LaunchColsBandIn>(int(p->Nos[1]), grid, int(p->Nos[0]), fg, p->twiddle[1].as<float2>(), -1.0f,
                                           s);
        GT_HIP(hipGetLastError());
    }
    if (f->n_c2nc)
        dispatch_fast<LaunchFusedC2NC>(f->KW, f->c2nc.as<gt::MfGridArgs>(), f->c2nc_units.as<gt::GridUnit>(),
                                       f->n_c2nc, s);
    GT_HIP(hipGetLastError());
}

void gt_host::fused_adjoint(FusedGridding* f, bool from_data, hipStream_t s) {
    const gt::MfGridArgs* args = (from_data ? f->nc2c_data : f->nc2c_int).as<gt::MfGridArgs>();
    if (f->n_nc2c) {
        auto k = f->dm == 0 ? gt::nc2c_mfma2d_batch_kernel<0, kFusedCW>
                            : (f->dm == 1 ? gt::nc2c_mfma2d_batch_kernel<1, kFusedCW>
                                          : gt::nc2c_mfma2d_batch_kernel<2, kFusedCW>);
        hipLaunchKernelGGL(k, dim3((unsigned)f->n_nc2c), dim3(64), 0, s, args, f->nc2c_units.as<gt::GridUnit>());
        GT_HIP(hipGetLastError());
    }
    if (f->n_red) {
        hipLaunchKernelGGL(gt::nc2c_reduce2d_batch_kernel, dim3((unsigned)f->n_red, (unsigned)f->red_parts), dim3(256),
                           0, s, args, f->reds.as<const gt::RedItem*>(), f->red_units.as<gt::GridUnit>());
        GT_HIP(hipGetLastError());
    }
    for (auto& g : f->groups) {
        gt_nufft_plan* p = g.rep;
        const float scale = centring_global_sign(p) / std::sqrt(float(p->G()));
        const gt::FftGeom fg = fft_geom(p, g.nb, false, scale);
        float2* grid = g.grid.as<float2>();
        dispatch_fft_len<LaunchColsBandOut>(int(p->Nos[1]), grid, int(p->Nos[0]), fg, p->twiddle[1].as<float2>(),
                                            1.0f, s);
        GT_HIP(hipGetLastError());
        dispatch_fft_len<LaunchRowsCrop>(int(p->Nos[0]), (const float2*)grid, g.coil.as<float2>(), fg,
                                         p->twiddle[0].as<float2>(), 1.0f, s);
        GT_HIP(hipGetLastError());
    }
}

void* gt_host::fused_coil_images(const FusedGridding* f, int i) {
    const auto& g = f->groups[size_t(f->group_of[size_t(i)])];
    return g.coil.as<float2>() + f->boff[size_t(i)] * g.rep->Nimg();
}

std::vector<uint64_t> gt_host::fused_generations(const FusedGridding* f) {
    std::vector<uint64_t> out;
    for (const auto& q : f->probs) out.push_back(q.plan->generation);
    return out;
}

// ---- fixed device scratch of gt_nufft_compute_noalloc -----------------------
// The oversampled grid and the split tiles' partial sums of one call live here
// instead of in the plan's grow-only workspaces.  Each GPU holds its own copy
// of the array from the library's first use on it until the process ends.  In
// host code `scratch` names a host-side stand-in, not device memory: its
// device address is looked up once per device (scratch_address) and only that
// address is handed to kernels.
constexpr size_t kScratchBytes = size_t(64) << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

std::mutex scratch_mu;
std::map<int, void*> scratch_ptrs;  // device -> address of that device's `scratch`

// Address of `scratch` on `device`, which must be the current device.  The
// first call on a device looks it up; later calls return the kept address.
void* scratch_address(int device) {
    std::lock_guard<std::mutex> lock(scratch_mu);
    auto it = scratch_ptrs.find(device);
    if (it != scratch_ptrs.end()) return it->second;
    void* ptr = nullptr;
    GT_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    GT_REQUIRE(ptr, GT_ERR_DEVICE, "hipGetSymbolAddress returned no address for the scratch array");
    scratch_ptrs[device] = ptr;
    return ptr;
}

// Stage timing draws HIP events from a pool that may grow; switched off for
// the duration of a no-allocation call.
struct ProfilingPaused {
    gt_nufft_plan* p;
    bool was;
    explicit ProfilingPaused(gt_nufft_plan* p_) : p(p_), was(p_->profiling) { p->profiling = false; }
    ~ProfilingPaused() { p->profiling = was; }
};

// gt_nufft_compute on the fused 2D path with the grid and partial tiles in
// `scratch`: nothing is allocated.  Everything is checked on the host first.
void compute_noalloc(gt_nufft_plan* p, const void* in, const size_t* in_dims, int in_ndim, void* out,
                     const size_t* out_dims, int out_ndim, const float* dcw, size_t dcw_len, int mode,
                     hipStream_t s) {
    GT_REQUIRE(in && out, GT_ERR_INVALID_ARGUMENT, "null input or output");
    GT_REQUIRE(mode >= 0 && mode <= 3, GT_ERR_INVALID_ARGUMENT, "unknown NFFT_comp_mode");
    GT_REQUIRE(!dcw || dcw_len > 0, GT_ERR_INVALID_ARGUMENT, "dcw given with zero length");
    GT_REQUIRE(!p->dbl && p->fast2d && p->fft_fused && fft2d_strategy() == 2, GT_ERR_UNSUPPORTED,
               "gt_nufft_compute_noalloc takes float 2D plans whose oversampled sizes are fused FFT lengths "
               "(96, 128, 144, 160, 192, 240, 256, 288, 320, 384)");
    GT_REQUIRE(p->conv_type == GT_CONV_STANDARD, GT_ERR_UNSUPPORTED,
               "gt_nufft_compute_noalloc takes STANDARD convolution only");
    const bool image_in = (mode == GT_FORWARDS_C2NC || mode == GT_BACKWARDS_C2NC);
    Shape sh = image_in ? image_shape(p, in_dims, in_ndim, true) : image_shape(p, out_dims, out_ndim, true);
    const size_t selems = image_in ? elements(out_dims, out_ndim) : elements(in_dims, in_ndim);
    check_batches(p, sh, selems);
    GT_REQUIRE(!sh.oversampled, GT_ERR_UNSUPPORTED,
               "gt_nufft_compute_noalloc takes an image of the matrix size, not the oversampled one");
    const size_t grid_bytes = sh.nb * p->G() * sizeof(float2);
    const size_t part_off = (grid_bytes + 255) & ~size_t(255);
    const size_t part_bytes =
        image_in ? 0 : size_t(p->bn.n_partials) * sh.C * size_t(p->bn.geom.tile_cells) * sizeof(float2);
    if (part_off + part_bytes > kScratchBytes) {
        std::ostringstream os;
        os << "gt_nufft_compute_noalloc needs " << part_off + part_bytes << " bytes of scratch (grid "
           << grid_bytes << ", partial tiles " << part_bytes << "), more than its fixed " << kScratchBytes
           << " bytes; use gt_nufft_compute";
        throw gt_host::gt_error(GT_ERR_UNSUPPORTED, os.str());
    }
    set_device(p);
    unsigned char* base = static_cast<unsigned char*>(scratch_address(p->device));
    ProfilingPaused paused(p);
    const float scale = centring_global_sign(p) / std::sqrt(float(p->G()));
    compute_fast2d_fused(p, in, out, sh, dcw, dcw_len, mode, scale, s, nullptr,
                         reinterpret_cast<float2*>(base), base + part_off);
}

}  // namespace

bool gt_host::plan_is_double(const gt_nufft_plan* p) { return p->dbl; }
uint64_t gt_host::plan_generation(const gt_nufft_plan* p) { return p->generation; }

bool gt_host::forward_with_coil_maps(gt_nufft_plan* p, const void* x, const void* csm, size_t C, void* out,
                                     const float* dcw, size_t dcw_len, hipStream_t s) {
    if (!(p->fast2d && p->fft_fused && fft2d_strategy() == 2)) return false;
    GT_REQUIRE(x && csm && out && C > 0, GT_ERR_INVALID_ARGUMENT, "null argument");
    GT_REQUIRE(p->M > 0 && p->F > 0, GT_ERR_NOT_PREPROCESSED, "plan has not been preprocessed");
    GT_REQUIRE(!dcw || dcw_len > 0, GT_ERR_INVALID_ARGUMENT, "dcw given with zero length");
    set_device(p);
    Shape sh;
    sh.oversampled = false;
    sh.nb = p->F * C;
    sh.C = C;
    const float scale = centring_global_sign(p) / std::sqrt(float(p->G()));
    compute_fast2d_fused(p, x, out, sh, dcw, dcw_len, GT_FORWARDS_C2NC, scale, s, static_cast<const float2*>(csm));
    return true;
}

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

const char* gt_last_error(void) { return gt_host::last_error().c_str(); }

gt_status gt_nufft_create(gt_nufft_plan** out, const gt_nufft_desc* d) {
    return guarded([&] {
        GT_REQUIRE(out && d, GT_ERR_INVALID_ARGUMENT, "null plan or descriptor");
        *out = nullptr;
        GT_REQUIRE(d->D >= 1 && d->D <= 4, GT_ERR_UNSUPPORTED, "D must be 1 to 4");
        GT_REQUIRE(d->precision == GT_FLOAT || d->precision == GT_DOUBLE, GT_ERR_UNSUPPORTED,
                   "precision must be GT_FLOAT or GT_DOUBLE");
        // NFFT<cuNDArray, double, D>::make_plan (cuNFFT.cu:219-227)
        GT_REQUIRE(d->precision == GT_FLOAT || d->conv_type == GT_CONV_STANDARD, GT_ERR_UNSUPPORTED,
                   "Only standard convolution type supported for doubles");
        GT_REQUIRE(d->conv_type == GT_CONV_STANDARD || d->conv_type == GT_CONV_ATOMIC ||
                       d->conv_type == GT_CONV_SPARSE_MATRIX, GT_ERR_INVALID_ARGUMENT,
                   "Invalid convolution type provided.");
        GT_REQUIRE(std::isfinite(d->W) && d->W >= 1.8, GT_ERR_INVALID_ARGUMENT,
                   "Kernel width must be equal to or larger than 1.8.");
        GT_REQUIRE(d->os_factor == 0.0 || d->os_factor >= 1.0, GT_ERR_INVALID_ARGUMENT,
                   "Oversampling factor must greater than or equal to 1.");
        auto p = std::make_unique<gt_nufft_plan>();
        p->track();
        p->D = d->D;
        p->W = d->W;
        p->Wf = float(d->W);
        p->radius = p->Wf / 2.0f;
        p->one_over_W = 1.0f / p->Wf;
        p->conv_type = d->conv_type;
        p->dbl = d->precision == GT_DOUBLE;
        p->radius_d = d->W / 2.0;
        p->one_over_W_d = 1.0 / d->W;
        for (int i = 0; i < p->D; i++) {
            GT_REQUIRE(d->N[i] >= 1, GT_ERR_INVALID_ARGUMENT, "matrix size must be positive");
            p->N[i] = d->N[i];
            float osf;
            double osf_d;
            if (d->os_factor != 0.0) {
                // NFFT.hpp:23-37: N_os = size_t(alpha * N) in REAL; KaiserKernel os ctor (ConvolutionKernel.hpp:241-254)
                const float a = float(d->os_factor);
                p->Nos[i] = p->dbl ? size_t(d->os_factor * double(d->N[i])) : size_t(a * float(d->N[i]));
                osf = a;
                osf_d = d->os_factor;
            } else {
                p->Nos[i] = d->N_os[i];
                osf = float(unsigned(d->N_os[i])) / float(unsigned(d->N[i]));
                osf_d = double(unsigned(d->N_os[i])) / double(unsigned(d->N[i]));
            }
            GT_REQUIRE(p->Nos[i] >= p->N[i], GT_ERR_INVALID_ARGUMENT,
                       "Oversampled matrix size must be greater than or equal to matrix size.");
            GT_REQUIRE(p->Nos[i] % 2 == 0, GT_ERR_UNSUPPORTED, "oversampled matrix size must be even");
            GT_REQUIRE(p->Nos[i] < (size_t(1) << 20), GT_ERR_UNSUPPORTED, "oversampled matrix size too large");
            GT_REQUIRE(double(p->Nos[i]) > d->W, GT_ERR_INVALID_ARGUMENT,
                       "kernel width must be smaller than the oversampled matrix size");
            p->beta[i] = gt::kb_beta(p->Wf, osf);
            p->beta_d[i] = gt::kb_beta_d(d->W, osf_d);
            p->nos_d[i] = double(p->Nos[i]);
            GT_REQUIRE(std::isfinite(p->beta[i]), GT_ERR_INVALID_ARGUMENT,
                       "Kaiser-Bessel beta is not finite for this width and oversampling factor");
        }
        p->KW = int(std::floor(2.0f * p->radius)) + 1;
        GT_REQUIRE(p->KW >= 2 && p->KW <= 9, GT_ERR_UNSUPPORTED, "kernel width must be below 9");
        int dev = d->device;
        if (dev < 0) GT_HIP(hipGetDevice(&dev));
        p->device = dev;
        set_device(p.get());
        GT_HIP(hipDeviceGetAttribute(&p->cus, hipDeviceAttributeMultiprocessorCount, dev));
        (void)scratch_address(dev);  // once per device, with the device current
        for (int i = 0; i < p->D; i++) {
            p->filt_host[i] = deapod_1d(p->Nos[i], p->radius, p->one_over_W, p->beta[i]);
            p->filt[i].reserve(p->Nos[i] * sizeof(float));
            GT_HIP(hipMemcpy(p->filt[i].ptr, p->filt_host[i].data(), p->Nos[i] * sizeof(float),
                             hipMemcpyHostToDevice));
            if (p->dbl) {
                const std::vector<double> fd = deapod_1d_d(p->Nos[i], p->radius_d, p->one_over_W_d, p->beta_d[i]);
                p->filt_d[i].reserve(p->Nos[i] * sizeof(double));
                GT_HIP(hipMemcpy(p->filt_d[i].ptr, fd.data(), p->Nos[i] * sizeof(double), hipMemcpyHostToDevice));
            }
        }
        p->flag.reserve(sizeof(int));
        setup_geometry(p.get());
        p->fft_fused = p->D == 2 && !p->dbl && fused_fft_len(int(p->Nos[0])) && fused_fft_len(int(p->Nos[1]));
        p->fft_fused3 = p->D == 3 && !p->dbl && fused_fft_len(int(p->Nos[0])) && fused_fft_len(int(p->Nos[1])) &&
                        fused_fft_len(int(p->Nos[2]));
        if (p->fft_fused || p->fft_fused3) {
            for (int i = 0; i < p->D; i++) {
                const size_t L = p->Nos[i];
                std::vector<float> tw(2 * L);
                for (size_t m = 0; m < L; m++) {
                    const double a = -2.0 * M_PI * double(m) / double(L);
                    tw[2 * m] = float(std::cos(a));
                    tw[2 * m + 1] = float(std::sin(a));
                }
                p->twiddle[i].reserve(L * sizeof(float2));
                GT_HIP(hipMemcpy(p->twiddle[i].ptr, tw.data(), L * sizeof(float2), hipMemcpyHostToDevice));
            }
        }
        *out = p.release();
    });
}

gt_status gt_nufft_destroy(gt_nufft_plan* p) {
    return guarded([&] {
        if (!p) return;
        (void)hipSetDevice(p->device);
        delete p;
    });
}

gt_status gt_nufft_set_tiling(gt_nufft_plan* p, const int* tile, int max_samples_per_item) {
    return guarded([&] {
        GT_REQUIRE(p, GT_ERR_INVALID_ARGUMENT, "null plan");
        for (int d = 0; d < 3; d++) {
            int t = tile ? tile[d] : 0;
            GT_REQUIRE(t <= 1024, GT_ERR_INVALID_ARGUMENT, "tile edge too large");
            p->tile_req[d] = t;
        }
        p->max_items_req = max_samples_per_item;
        p->generation++;
        setup_geometry(p);
        int cells = p->geom.tile_cells;
        GT_REQUIRE(size_t(2) * 4 * cells <= 65536, GT_ERR_INVALID_ARGUMENT, "tile does not fit in LDS");
        p->prep_c2nc = p->prep_nc2c = false;  // bins depend on the tiles: preprocess again
        p->M = p->F = 0;
    });
}

gt_status gt_nufft_preprocess(gt_nufft_plan* p, const void* traj, size_t M, size_t F, int mode,
                              gt_stream_t stream) {
    return guarded([&] {
        GT_REQUIRE(p, GT_ERR_INVALID_ARGUMENT, "null plan");
        GT_REQUIRE(traj, GT_ERR_INVALID_ARGUMENT, "cannot preprocess a null trajectory");
        GT_REQUIRE(M > 0 && F > 0, GT_ERR_INVALID_ARGUMENT, "trajectory must have samples and frames");
        GT_REQUIRE(mode >= GT_PREP_C2NC && mode <= GT_PREP_ALL, GT_ERR_INVALID_ARGUMENT,
                   "unknown NFFT_prep_mode");
        GT_REQUIRE(M * F < (size_t(1) << 31), GT_ERR_UNSUPPORTED, "too many samples");
        hipStream_t s = S(stream);
        set_device(p);
        p->generation++;
        p->prep_c2nc = p->prep_nc2c = false;
        p->M = p->F = 0;
        const long long n = (long long)(M * F * p->D);
        p->traj.reserve(size_t(n) * sizeof(float));
        GT_HIP(hipMemsetAsync(p->flag.ptr, 0, sizeof(int), s));
        if (p->dbl) {
            // traj: double D-vectors; positions in double, a float copy for the binning
            p->traj_d.reserve(size_t(n) * sizeof(double));
            hipLaunchKernelGGL(gt::scale_traj_d_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s,
                               (const double*)traj, p->traj_d.as<double>(), p->traj.as<float>(), n, p->D, p->geom,
                               p->flag.as<int>());
        } else {
            hipLaunchKernelGGL(gt::scale_traj_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s,
                               (const float*)traj, p->traj.as<float>(), n, p->D, p->geom, p->flag.as<int>());
        }
        GT_HIP(hipGetLastError());
        int bad = 0;
        GT_HIP(hipMemcpyAsync(&bad, p->flag.ptr, sizeof(int), hipMemcpyDeviceToHost, s));
        GT_HIP(hipStreamSynchronize(s));
        GT_REQUIRE(!bad, GT_ERR_RANGE, "trajectory out of range [-1/2, 1/2] (or not finite)");
        p->M = M;
        p->F = F;
        if (mode == GT_PREP_C2NC || mode == GT_PREP_ALL) {
            build_c2nc(p, s);
            p->prep_c2nc = true;
        }
        if (mode == GT_PREP_NC2C || mode == GT_PREP_ALL) {
            if (p->conv_type != GT_CONV_ATOMIC) build_nc2c(p, s);
            if (p->conv_type == GT_CONV_SPARSE_MATRIX) {
                const long long FM = (long long)(M * F);
                p->row_inv.reserve(size_t(FM) * sizeof(float));
                hipLaunchKernelGGL(gt::row_norm_kernel, dim3((unsigned)((FM + 255) / 256)), dim3(256), 0, s,
                                   p->traj.as<float>(), p->row_inv.as<float>(), FM, p->geom);
                GT_HIP(hipGetLastError());
            }
            p->prep_nc2c = true;
        }
    });
}

gt_status gt_nufft_compute(gt_nufft_plan* p, const void* in, const size_t* in_dims, int in_ndim, void* out,
                           const size_t* out_dims, int out_ndim, const float* dcw, size_t dcw_len,
                           int comp_mode, gt_stream_t stream) {
    return guarded([&] {
        GT_REQUIRE(p, GT_ERR_INVALID_ARGUMENT, "null plan");
        compute_impl(p, in, in_dims, in_ndim, out, out_dims, out_ndim, dcw, dcw_len, comp_mode, S(stream));
    });
}

gt_status gt_nufft_compute_noalloc(gt_nufft_plan* p, const void* in, const size_t* in_dims, int in_ndim, void* out,
                                   const size_t* out_dims, int out_ndim, const float* dcw, size_t dcw_len,
                                   int comp_mode, gt_stream_t stream) {
    return guarded([&] {
        GT_REQUIRE(p, GT_ERR_INVALID_ARGUMENT, "null plan");
        compute_noalloc(p, in, in_dims, in_ndim, out, out_dims, out_ndim, dcw, dcw_len, comp_mode, S(stream));
    });
}

gt_status gt_nufft_mult_MH_M(gt_nufft_plan* p, const void* in, void* out, const size_t* dims, int ndim,
                             const float* dcw, size_t dcw_len, gt_stream_t stream) {
    return guarded([&] {
        GT_REQUIRE(p, GT_ERR_INVALID_ARGUMENT, "null plan");
        GT_REQUIRE(p->M > 0, GT_ERR_NOT_PREPROCESSED, "plan has not been preprocessed");
        Shape sh = image_shape(p, dims, ndim, true);
        const size_t nb = sh.nb;
        p->tmp_samples.reserve(std::max<size_t>(nb * p->M, 1) * (p->dbl ? sizeof(double2) : sizeof(float2)));
        const size_t sd[1] = {nb * p->M};
        compute_impl(p, in, dims, ndim, p->tmp_samples.ptr, sd, 1, dcw, dcw_len, GT_FORWARDS_C2NC, S(stream));
        compute_impl(p, p->tmp_samples.ptr, sd, 1, out, dims, ndim, dcw, dcw_len, GT_BACKWARDS_NC2C, S(stream));
    });
}

gt_status gt_nufft_convolve(gt_nufft_plan* p, const void* in, const size_t* in_dims, int in_ndim, void* out,
                            const size_t* out_dims, int out_ndim, int conv_mode, int accumulate,
                            gt_stream_t stream) {
    return guarded([&] {
        GT_REQUIRE(p && in && out, GT_ERR_INVALID_ARGUMENT, "null plan, input or output");
        GT_REQUIRE(conv_mode == GT_CONV_C2NC || conv_mode == GT_CONV_NC2C, GT_ERR_INVALID_ARGUMENT,
                   "unknown NFFT_conv_mode");
        const bool c2nc = conv_mode == GT_CONV_C2NC;
        Shape sh = c2nc ? image_shape(p, in_dims, in_ndim, false) : image_shape(p, out_dims, out_ndim, false);
        check_batches(p, sh, c2nc ? elements(out_dims, out_ndim) : elements(in_dims, in_ndim));
        set_device(p);
        const gt::Layout L = caller_layout(p);
        if (p->dbl) {
            if (c2nc) conv_c2nc_d(p, in, out, sh.C, 0, nullptr, 0, accumulate != 0, S(stream));
            else conv_nc2c_d(p, in, out, sh.C, 0, nullptr, 0, accumulate != 0, S(stream));
        } else if (c2nc) {
            conv_c2nc(p, in, L, out, sh.C, 0, nullptr, 0, accumulate != 0, S(stream));
        } else {
            conv_nc2c(p, in, out, L, sh.C, 0, nullptr, 0, accumulate != 0, S(stream));
        }
    });
}

gt_status gt_nufft_fft(gt_nufft_plan* p, void* data, const size_t* dims, int ndim, int fft_mode, int do_scale,
                       gt_stream_t stream) {
    return guarded([&] {
        GT_REQUIRE(p && data, GT_ERR_INVALID_ARGUMENT, "null plan or data");
        GT_REQUIRE(fft_mode == GT_FFT_FORWARDS || fft_mode == GT_FFT_BACKWARDS, GT_ERR_INVALID_ARGUMENT,
                   "unknown NFFT_fft_mode");
        Shape sh = image_shape(p, dims, ndim, false);
        set_device(p);
        hipStream_t s = S(stream);
        const long long total = (long long)(sh.nb * p->G());
        if (p->dbl) {
            const double gsd = centring_global_sign(p);
            const double sd = do_scale ? gsd / std::sqrt(double(p->G())) : gsd;
            hipLaunchKernelGGL(gt::sign_scale_d_kernel, dim3(grid_blocks(total)), dim3(256), 0, s, (double2*)data,
                               (long long)sh.nb, p->geom, 1.0, 1);
            GT_HIP(hipGetLastError());
            run_fft(p, data, sh.nb, fft_mode == GT_FFT_FORWARDS, s);
            hipLaunchKernelGGL(gt::sign_scale_d_kernel, dim3(grid_blocks(total)), dim3(256), 0, s, (double2*)data,
                               (long long)sh.nb, p->geom, sd, 1);
            GT_HIP(hipGetLastError());
            return;
        }
        const float gs = centring_global_sign(p);
        const float scale = do_scale ? gs / std::sqrt(float(p->G())) : gs;
        // centred FFT = (-1)^(sum N_os/2) * T . FFT . T  (cuNDFFT.cpp:124-138 timeswitch pair)
        hipLaunchKernelGGL(gt::sign_scale_kernel, dim3(grid_blocks(total)), dim3(256), 0, s, (float2*)data,
                           (long long)sh.nb, p->geom, 1.0f, 1);
        GT_HIP(hipGetLastError());
        run_fft(p, data, sh.nb, fft_mode == GT_FFT_FORWARDS, s);
        hipLaunchKernelGGL(gt::sign_scale_kernel, dim3(grid_blocks(total)), dim3(256), 0, s, (float2*)data,
                           (long long)sh.nb, p->geom, scale, 1);
        GT_HIP(hipGetLastError());
    });
}

gt_status gt_nufft_deapodize(gt_nufft_plan* p, void* image, const size_t* dims, int ndim, int fourier_domain,
                             gt_stream_t stream) {
    return guarded([&] {
        GT_REQUIRE(p && image, GT_ERR_INVALID_ARGUMENT, "null plan or image");
        (void)fourier_domain;  // the filter is real and even: both domains use the same values
        Shape sh;
        try {
            sh = image_shape(p, dims, ndim, false);
        } catch (const gt_error&) {
            throw gt_error(GT_ERR_SHAPE, "Error: cuNFFT_impl::deapodize: ERROR: oversampled image not provided as input.");
        }
        set_device(p);
        if (p->dbl) {
            crop_deapod_d(p, image, image, sh.nb, true, 1.0, 0, S(stream));  // in place, as below
            return;
        }
        gt::PadCrop pc = padcrop(p, false, true, 1.0f, 0);
        const long long rows = (long long)(sh.nb * p->G() / p->Nos[0]);  // in place: each thread rewrites what it read
        hipLaunchKernelGGL(gt::crop_deapod_kernel, dim3((unsigned)((rows + gt::kRowsPerBlock - 1) / gt::kRowsPerBlock)),
                           dim3(64, gt::kRowsPerBlock), 0, S(stream), (const float2*)image, (float2*)image,
                           (long long)sh.nb, p->geom, pc);
        GT_HIP(hipGetLastError());
    });
}

gt_status gt_nufft_grad_traj(gt_nufft_plan* p, const void* image, const size_t* dims, int ndim,
                             const void* grad_samples, const size_t* sdims, int sndim, const float* dcw,
                             size_t dcw_len, float* grad_traj, gt_stream_t stream) {
    return guarded([&] {
        GT_REQUIRE(p && image && grad_samples && grad_traj, GT_ERR_INVALID_ARGUMENT, "null argument");
        GT_REQUIRE(!p->dbl, GT_ERR_UNSUPPORTED, "the trajectory gradient is provided for float plans");
        GT_REQUIRE(p->D <= 3, GT_ERR_UNSUPPORTED, "the trajectory gradient is provided for D <= 3");
        GT_REQUIRE(!dcw || dcw_len > 0, GT_ERR_INVALID_ARGUMENT, "dcw given with zero length");
        GT_REQUIRE(p->M > 0 && p->F > 0, GT_ERR_NOT_PREPROCESSED, "plan has not been preprocessed");
        Shape sh = image_shape(p, dims, ndim, true);
        check_batches(p, sh, elements(sdims, sndim));
        set_device(p);
        hipStream_t s = S(stream);
        // the grid of FORWARDS_C2NC, recomputed from the image (not stored by
        // the forward pass): pad, deapodize, centred FFT, caller layout
        const size_t G = p->G();
        const float scale = centring_global_sign(p) / std::sqrt(float(G));
        p->grid.reserve(sh.nb * G * sizeof(float2));
        pad_deapod(p, image, p->grid.ptr, sh.nb, sh.oversampled, scale, s);
        run_fft(p, p->grid.ptr, sh.nb, true, s);
        const long long dl = (long long)std::max<size_t>(dcw_len, 1);
        dispatch<LaunchTrajGrad>(p->D, p->KW, (const float2*)p->grid.ptr, (const float2*)grad_samples,
                                 p->traj.as<float>(), int(p->M), int(p->F), int(sh.C), p->geom, dcw, dl, grad_traj, s);
        GT_HIP(hipGetLastError());
    });
}

gt_status gt_nufft_get_matrix_size(const gt_nufft_plan* p, size_t* N) {
    return guarded([&] {
        GT_REQUIRE(p && N, GT_ERR_INVALID_ARGUMENT, "null argument");
        for (int d = 0; d < p->D; d++) N[d] = p->N[d];
    });
}

gt_status gt_nufft_get_matrix_size_os(const gt_nufft_plan* p, size_t* N_os) {
    return guarded([&] {
        GT_REQUIRE(p && N_os, GT_ERR_INVALID_ARGUMENT, "null argument");
        for (int d = 0; d < p->D; d++) N_os[d] = p->Nos[d];
    });
}

gt_status gt_nufft_get_W(const gt_nufft_plan* p, double* W) {
    return guarded([&] {
        GT_REQUIRE(p && W, GT_ERR_INVALID_ARGUMENT, "null argument");
        *W = p->W;
    });
}

gt_status gt_nufft_get_num_samples_frames(const gt_nufft_plan* p, size_t* M, size_t* F) {
    return guarded([&] {
        GT_REQUIRE(p && M && F, GT_ERR_INVALID_ARGUMENT, "null argument");
        *M = p->M;
        *F = p->F;
    });
}

gt_status gt_nufft_profile(gt_nufft_plan* p, int enable) {
    return gt_nufft_profile_stages(p, enable ? (1u << GT_NUM_STAGES) - 1u : 0u);
}

gt_status gt_nufft_profile_stages(gt_nufft_plan* p, unsigned stage_mask) {
    return guarded([&] {
        GT_REQUIRE(p, GT_ERR_INVALID_ARGUMENT, "null plan");
        p->stage_mask = stage_mask & ((1u << GT_NUM_STAGES) - 1u);
        p->profiling = p->stage_mask != 0;
        if (p->profiling) {
            // events come from the pool while recording: fill it now, so that a
            // timed loop records events without creating any
            set_device(p);
            while (p->event_pool.size() < kEventPoolSize) {
                hipEvent_t e = nullptr;
                GT_HIP(hipEventCreate(&e));
                p->event_pool.push_back(e);
            }
        }
    });
}

gt_status gt_nufft_profile_read(gt_nufft_plan* p, double* total_ms, int* counts, int n) {
    return guarded([&] {
        GT_REQUIRE(p && total_ms && counts && n > 0, GT_ERR_INVALID_ARGUMENT, "null argument");
        for (int i = 0; i < n; i++) {
            total_ms[i] = 0.0;
            counts[i] = 0;
        }
        set_device(p);
        std::vector<gt_nufft_plan::Record> recs;
        recs.swap(p->records);
        for (auto& r : recs) {
            float ms = 0.0f;
            GT_HIP(hipEventSynchronize(r.b));
            GT_HIP(hipEventElapsedTime(&ms, r.a, r.b));
            if (r.stage < n) {
                total_ms[r.stage] += ms;
                counts[r.stage] += 1;
            }
            p->event_pool.push_back(r.a);
            p->event_pool.push_back(r.b);
        }
    });
}

// ---- device plumbing --------------------------------------------------------
gt_status gt_device_count(int* count) {
    return guarded([&] {
        GT_REQUIRE(count, GT_ERR_INVALID_ARGUMENT, "null argument");
        int n = 0;
        *count = 0;
        const hipError_t e = hipGetDeviceCount(&n);
        if (e == hipErrorNoDevice) return;  // a host without GPUs: count 0
        GT_HIP(e);
        *count = n;
    });
}
gt_status gt_set_device(int device) { return guarded([&] { GT_HIP(hipSetDevice(device)); }); }
gt_status gt_malloc(void** ptr, size_t bytes) {
    return guarded([&] {
        GT_REQUIRE(ptr, GT_ERR_INVALID_ARGUMENT, "null argument");
        GT_HIP(hipMalloc(ptr, std::max<size_t>(bytes, 1)));
    });
}
gt_status gt_free(void* ptr) { return guarded([&] { if (ptr) GT_HIP(hipFree(ptr)); }); }
gt_status gt_memcpy_htod(void* dst, const void* src, size_t bytes, gt_stream_t s) {
    return guarded([&] {
        GT_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, S(s)));
        GT_HIP(hipStreamSynchronize(S(s)));
    });
}
gt_status gt_memcpy_dtoh(void* dst, const void* src, size_t bytes, gt_stream_t s) {
    return guarded([&] {
        GT_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, S(s)));
        GT_HIP(hipStreamSynchronize(S(s)));
    });
}
gt_status gt_memcpy_dtod(void* dst, const void* src, size_t bytes, gt_stream_t s) {
    return guarded([&] { GT_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, S(s))); });
}
gt_status gt_memset(void* ptr, int value, size_t bytes, gt_stream_t s) {
    return guarded([&] { GT_HIP(hipMemsetAsync(ptr, value, bytes, S(s))); });
}
gt_status gt_array_add(void* y, const void* x, size_t n, int precision, gt_stream_t s) {
    return guarded([&] {
        GT_REQUIRE(precision == GT_FLOAT || precision == GT_DOUBLE, GT_ERR_INVALID_ARGUMENT,
                   "precision must be GT_FLOAT or GT_DOUBLE");
        if (n == 0) return;
        GT_REQUIRE(y && x, GT_ERR_INVALID_ARGUMENT, "null array");
        const unsigned blocks = (unsigned)std::min<size_t>((n + 255) / 256, size_t(1) << 20);
        if (precision == GT_FLOAT)
            hipLaunchKernelGGL(gt::add_kernel<float>, dim3(blocks), dim3(256), 0, S(s), (float*)y, (const float*)x,
                               (long long)n);
        else
            hipLaunchKernelGGL(gt::add_kernel<double>, dim3(blocks), dim3(256), 0, S(s), (double*)y,
                               (const double*)x, (long long)n);
        GT_HIP(hipGetLastError());
    });
}
gt_status gt_array_sqrt(float* y, size_t n, gt_stream_t s) {
    return guarded([&] {
        if (n == 0) return;
        GT_REQUIRE(y, GT_ERR_INVALID_ARGUMENT, "null array");
        const unsigned blocks = (unsigned)std::min<size_t>((n + 255) / 256, size_t(1) << 20);
        hipLaunchKernelGGL(gt::sqrt_kernel, dim3(blocks), dim3(256), 0, S(s), y, (long long)n);
        GT_HIP(hipGetLastError());
    });
}
gt_status gt_array_scale(float* y, size_t n, float a, gt_stream_t s) {
    return guarded([&] {
        if (n == 0) return;
        GT_REQUIRE(y, GT_ERR_INVALID_ARGUMENT, "null array");
        const unsigned blocks = (unsigned)std::min<size_t>((n + 255) / 256, size_t(1) << 20);
        hipLaunchKernelGGL(gt::scale_kernel, dim3(blocks), dim3(256), 0, S(s), y, (long long)n, a);
        GT_HIP(hipGetLastError());
    });
}
gt_status gt_array_mul_weights(void* x, const float* w, size_t n, size_t batches, gt_stream_t s) {
    return guarded([&] {
        if (n == 0 || batches == 0) return;
        GT_REQUIRE(x && w, GT_ERR_INVALID_ARGUMENT, "null array");
        GT_REQUIRE(n <= size_t(1) << 40 && batches <= (size_t(1) << 40) / n, GT_ERR_INVALID_ARGUMENT,
                   "array too large");
        const size_t total = n * batches;
        const unsigned blocks = (unsigned)std::min<size_t>((total + 255) / 256, size_t(1) << 20);
        hipLaunchKernelGGL(gt::mul_weights_kernel, dim3(blocks), dim3(256), 0, S(s), (float2*)x, w, (long long)n,
                           (long long)total);
        GT_HIP(hipGetLastError());
    });
}
gt_status gt_stream_synchronize(gt_stream_t s) { return guarded([&] { GT_HIP(hipStreamSynchronize(S(s))); }); }
gt_status gt_device_synchronize(void) { return guarded([&] { GT_HIP(hipDeviceSynchronize()); }); }

}  // extern "C"
