// Internal definitions shared by the tbplas_amd HIP sources.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <memory>
#include <utility>
#include <string>
#include <vector>

#include "../../include/tbplas_amd.h"

namespace tbg {

// ---------------------------------------------------------------------------
// Errors: every entry point returns a status code; the message of the last
// failure on the calling thread is kept for tbg_last_error().
// ---------------------------------------------------------------------------
void set_error(const std::string& msg);

struct Error {
    int code;
    std::string msg;
};

#define TBG_HIP(call)                                                          \
    do {                                                                       \
        hipError_t e_ = (call);                                                \
        if (e_ != hipSuccess)                                                  \
            throw ::tbg::Error{TBG_EHIP, std::string(#call) + ": " +           \
                                             hipGetErrorString(e_)};           \
    } while (0)

#define TBG_REQUIRE(cond, msg)                                                 \
    do {                                                                       \
        if (!(cond))                                                           \
            throw ::tbg::Error{TBG_EINVAL, (msg)};                             \
    } while (0)

// Device memory owned by the library (freed in the destructor), or borrowed
// from memory the library does not allocate (never freed here).
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    bool owned = true;
    DevBuf() = default;
    explicit DevBuf(size_t b) { alloc(b); }
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept : p(o.p), bytes(o.bytes), owned(o.owned) {
        o.p = nullptr;
        o.bytes = 0;
        o.owned = true;
    }
    DevBuf& operator=(DevBuf&& o) noexcept {
        if (this != &o) {
            release();
            p = o.p;
            bytes = o.bytes;
            owned = o.owned;
            o.p = nullptr;
            o.bytes = 0;
            o.owned = true;
        }
        return *this;
    }
    ~DevBuf() { release(); }
    void alloc(size_t b) {
        release();
        if (b == 0) b = 16;
        hipError_t e = hipMalloc(&p, b);
        if (e != hipSuccess) {
            p = nullptr;
            throw Error{TBG_EOOM, "hipMalloc(" + std::to_string(b) + " bytes) failed: " +
                                      hipGetErrorString(e)};
        }
        bytes = b;
        owned = true;
    }
    // use b bytes at q (device memory the caller keeps alive)
    void borrow(void* q, size_t b) {
        release();
        p = q;
        bytes = b;
        owned = false;
    }
    void ensure(size_t b) { if (b > bytes) alloc(b); }
    void release() {
        if (p && owned) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
        owned = true;
    }
    template <typename T> T* as() const { return reinterpret_cast<T*>(p); }
};

// Complex storage type of a working precision.
template <typename R> struct CplxT;
template <> struct CplxT<float> { using T = float2; };
template <> struct CplxT<double> { using T = double2; };

// Rows per slice of the sliced-ELL layout: one wavefront owns one slice.
constexpr int kSlice = 64;
constexpr int kWave = 64;

struct Blocking;

}  // namespace tbg

// Library handle: one per host thread; all work is ordered on `stream`.
struct tbg_handle {
    int device = 0;
    hipStream_t stream = nullptr;
    int num_cus = 256;
    int use_graph = 1;
    // Launch shape of the term kernels (set by tbg_set_tuning / the tuner):
    // persistent blocks per CU and LDS staging of the slice's own rows.
    int blocks_per_cu = 16;
    int stage = 1;
    int real_layout = 1;      // store real values when H has no imaginary part
    int blocking = 3;         // terms per pass where H has spatial tiles (3 to 5; 0: one)
    int block_rows = 0;       // cap on rows owned per workgroup (0: from LDS)
    int max_lds = 65536;      // LDS bytes one workgroup may use
    tbg::DevBuf scratch;      // reduction partials and small tables
    tbg::DevBuf counter;      // step counter for graph-replayed reductions
    // three-term tables of recent batches (key: state row bytes, limits and
    // the instances' ids); a null entry records tiles that do not fit
    std::vector<std::pair<std::string, std::shared_ptr<tbg::Blocking>>> batch_blk;
    // Per-launch timing (tbg_set_kernel_timing): a pair of events around
    // every propagation and correlation launch, tagged with the launch kind
    // (tbg_kernel_times sums them); events are reused from ev_pool.
    int timing = 0;
    int timing_hold = 0;      // > 0 inside one-off set-up work (not timed)
    struct Timed {
        int kind;
        hipEvent_t e0, e1;
    };
    std::vector<Timed> timed;
    std::vector<hipEvent_t> ev_pool;
};

namespace tbg {

// Kinds of timed launches: the one-term kernel, the AC / DC correlation
// kernels, the Hall moments' Gram block, and the tile pass as kTimePass +
// 10 * mode + terms.
enum TimedKind : int { kTimeTerm = 1, kTimeAcCorr = 2, kTimeQe = 3, kTimeDcDot = 4,
                       kTimeGram = 5, kTimePass = 100 };

// Runs launch() on h->stream, bracketed by two events when timing is on.
template <typename F>
inline void timed_launch(tbg_handle* h, int kind, F&& launch) {
    if (!h->timing || h->timing_hold > 0) {
        launch();
        return;
    }
    auto take = [&]() {
        hipEvent_t e;
        if (!h->ev_pool.empty()) {
            e = h->ev_pool.back();
            h->ev_pool.pop_back();
        } else {
            TBG_HIP(hipEventCreate(&e));
        }
        return e;
    };
    TBG_REQUIRE(h->timed.size() < (1u << 16),
                "too many timed launches pending: read them with tbg_kernel_times");
    hipEvent_t e0 = take(), e1 = take();
    h->timed.push_back({kind, e0, e1});
    TBG_HIP(hipEventRecord(e0, h->stream));
    launch();
    TBG_HIP(hipEventRecord(e1, h->stream));
}

// Launches inside its scope are not timed (one-off set-up of a run, such as
// applying the Fermi operators before the time loop).
struct TimingHold {
    tbg_handle* h;
    explicit TimingHold(tbg_handle* hh) : h(hh) { ++h->timing_hold; }
    ~TimingHold() { --h->timing_hold; }
    TimingHold(const TimingHold&) = delete;
    TimingHold& operator=(const TimingHold&) = delete;
};

}  // namespace tbg

// Hamiltonian prepared for propagation: sliced-ELL (64-row slices, column
// major inside a slice), int32 columns, complex values in the working
// precision; optional current operators on the same structure.
struct tbg_ham {
    tbg_handle* h = nullptr;
    int64_t n = 0;
    int64_t nnz = 0;
    int64_t nslices = 0;
    int64_t padded = 0;       // stored elements including padding
    int precision = TBG_FP64;
    int max_width = 0;
    tbg::DevBuf slice_off;    // int64 [nslices+1] element offsets
    tbg::DevBuf slice_w;      // int32 [nslices]
    tbg::DevBuf cols;         // int32 [padded]
    tbg::DevBuf vals;         // complex<R> [padded], or R [padded] when real_vals
    bool real_vals = false;   // every imaginary part of H is exactly zero
    tbg::DevBuf jx, jy;       // complex<R> [padded] current operators (optional)
    bool has_current = false;
    // Optional row order: layout row i holds the caller's row perm[i]
    // (int32 [n]); states are packed and unpacked through it.
    tbg::DevBuf perm;
    bool ordered = false;
    const int* perm_ptr() const { return ordered ? perm.as<int>() : nullptr; }
    // Spatial tiles of the layout order (host, ntiles + 1 row offsets) and
    // the temporal-blocking tables built from them on first use.
    std::vector<int64_t> tiles;
    mutable std::shared_ptr<tbg::Blocking> blk;
    mutable bool blk_failed = false;
    mutable int blk_row_bytes = 0;
    mutable int blk_rows_cap = 0;
    mutable int blk_depth = 0;
    uint64_t uid = 0;         // unique per Hamiltonian created in the process
};
