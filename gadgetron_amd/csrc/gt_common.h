// gt_common.h -- host-side plumbing shared by the translation units of
// libgtnufft: status exceptions, the guarded C-ABI wrapper, grow-only device
// buffers.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <new>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/gt_nufft.h"

namespace gt_host {

// Thread-local message behind gt_last_error() (defined once, in gt_plan.hip).
std::string& last_error();

struct gt_error : std::runtime_error {
    gt_status status;
    gt_error(gt_status s, const std::string& m) : std::runtime_error(m), status(s) {}
};

#define GT_HIP(call)                                                                      \
    do {                                                                                  \
        hipError_t e_ = (call);                                                           \
        if (e_ != hipSuccess) {                                                           \
            std::ostringstream os_;                                                       \
            os_ << #call << ": " << hipGetErrorString(e_);                                \
            throw ::gt_host::gt_error(e_ == hipErrorOutOfMemory ? GT_ERR_OUT_OF_MEMORY : GT_ERR_DEVICE, \
                                      os_.str());                                         \
        }                                                                                 \
    } while (0)

#define GT_REQUIRE(cond, status, msg)                             \
    do {                                                          \
        if (!(cond)) throw ::gt_host::gt_error((status), (msg)); \
    } while (0)

// Runs f, turning exceptions into a status and the thread-local message.
template <class F>
gt_status guarded(F&& f) {
    try {
        f();
        return GT_OK;
    } catch (const gt_error& e) {
        last_error() = e.what();
        return e.status;
    } catch (const std::bad_alloc&) {
        last_error() = "host allocation failed";
        return GT_ERR_OUT_OF_MEMORY;
    } catch (const std::exception& e) {
        last_error() = e.what();
        return GT_ERR_INTERNAL;
    }
}

inline hipStream_t S(gt_stream_t s) { return reinterpret_cast<hipStream_t>(s); }

// Grow-only device buffer.  `epoch`, when set, is the owner's generation
// counter: every reallocation bumps it, so work recorded against the old
// address (a captured hipGraph) can tell that it is stale.
struct DevBuf {
    void* ptr = nullptr;
    size_t bytes = 0;
    uint64_t* epoch = nullptr;
    void reserve(size_t n) {
        if (n <= bytes) return;
        if (epoch) ++*epoch;
        if (ptr) GT_HIP(hipFree(ptr));
        ptr = nullptr;
        bytes = 0;
        GT_HIP(hipMalloc(&ptr, n));
        bytes = n;
    }
    void release() {
        if (ptr) (void)hipFree(ptr);
        ptr = nullptr;
        bytes = 0;
    }
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept : ptr(o.ptr), bytes(o.bytes), epoch(o.epoch) {
        o.ptr = nullptr;
        o.bytes = 0;
    }
    ~DevBuf() { release(); }
    template <class T>
    T* as() const { return static_cast<T*>(ptr); }
};

// Plan-internal entry for the SENSE operator (gt_cg.hip): FORWARDS_C2NC of the
// image x [N..., F] weighted by the coil maps csm [N..., C] into samples
// [M, F, C], with the coil multiply done inside the first FFT pass.  Returns
// false, having done nothing, where the plan's compute has no such pass; the
// caller then multiplies the coil maps itself.
bool forward_with_coil_maps(gt_nufft_plan* p, const void* x, const void* csm, size_t C, void* out,
                            const float* dcw, size_t dcw_len, hipStream_t s);
// True for a GT_DOUBLE plan (the SENSE operator and CG solver take float plans).
bool plan_is_double(const gt_nufft_plan* p);
// The plan's generation: it changes whenever preprocess, set_tiling or a
// workspace reallocation moves or rewrites state that launched kernels read.
uint64_t plan_generation(const gt_nufft_plan* p);

// ---- fused gridding of many 2D problems (the fused batched CG-SENSE) --------
// Problems are grouped by (N, N_os).  Each group keeps one oversampled grid
// [y][its batches][x] and one coil-image buffer [batch][N1][N0]; a stage of
// every problem is one launch (gridding, split-tile reduction) or one launch
// per group (the FFT passes), from per-problem argument tables built here.
struct FusedProblem {
    gt_nufft_plan* plan;   // preprocessed float 2D plan (not owned)
    size_t C;              // coils; batches F C
    const float* dcw;      // NULL or per-sample weights
    size_t dcw_len;
    const void* csm;       // [N..., C] complex64
    const void* image;     // image the forward transform reads, [N..., F]
    const void* data;      // samples the first adjoint reads, [M, F, C]
};
struct FusedGridding;
// Why a problem cannot take part, or an empty string.
std::string fused_unsupported(const FusedProblem& p);
FusedGridding* fused_create(const std::vector<FusedProblem>& probs, hipStream_t s);
void fused_destroy(FusedGridding* f);
// E^H before the coil sum for every problem: samples (the callers' data, or
// the internal samples of the last fused_forward) -> coil images.
void fused_adjoint(FusedGridding* f, bool from_data, hipStream_t s);
// E for every problem: image (coil maps applied in the first FFT pass) ->
// internal samples.
void fused_forward(FusedGridding* f, hipStream_t s);
// Coil images of problem i after fused_adjoint, [N..., F, C] complex64.
void* fused_coil_images(const FusedGridding* f, int i);
// Generations of the bound plans, in problem order.
std::vector<uint64_t> fused_generations(const FusedGridding* f);
// Rebuilds a preprocessed 2D plan's NC2C work items with `len` table entries
// per item (0: the plan's own rule); the gridding result changes only in the
// summation order of split tiles.
void plan_retile_nc2c(gt_nufft_plan* p, int len, hipStream_t s);

}  // namespace gt_host
