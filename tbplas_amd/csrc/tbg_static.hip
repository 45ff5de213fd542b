// DOS correlation without device allocation (tbg_corr_dos_noalloc).
//
// Every device buffer of the run (the recurrence vectors, the packed initial
// states, the per-chunk partial sums, the correlation rows and the step
// counter) is carved from `scratch`, a file-scope device array of
// kScratchBytes that each GPU holds for the life of the process.  Its device
// address is looked up once per device when a handle is created
// (static_scratch_init) and only that address is passed to kernels and
// copies: in host code the array's name is a host-side stand-in, not a
// device address.
//
// The run is the one-term-per-pass schedule of tbg_corr_dos with the same
// chunk plan, so its correlations are bit-identical to tbg_corr_dos on a
// Hamiltonian without spatial tiles.  Calls on one GPU share the array:
// calls that overlap in time on one GPU (several handles or threads) race.
#include <cmath>
#include <mutex>

#include "tbg_engine.h"

namespace tbg {

constexpr size_t kScratchBytes = size_t(64) << 20;
constexpr int kScratchDevices = 64;

}  // namespace tbg

static __device__ __align__(256) unsigned char scratch[tbg::kScratchBytes];

namespace tbg {

static std::mutex g_scratch_mu;
static void* g_scratch_ptr[kScratchDevices] = {};

// Device address of `scratch` on `device` (which must be current): looked up
// once per device, at handle creation.
void static_scratch_init(int device) {
    TBG_REQUIRE(device >= 0 && device < kScratchDevices, "device index out of range");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[device] != nullptr) return;
    void* p = nullptr;
    TBG_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    TBG_REQUIRE(p != nullptr, "no device address for the scratch array");
    g_scratch_ptr[device] = p;
}

static void* scratch_ptr(int device) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    return (device >= 0 && device < kScratchDevices) ? g_scratch_ptr[device] : nullptr;
}

// chunk[c] = c * per (c < nchunks), chunk[nchunks] = ns; inst = {0, nchunks}:
// the plan make_plan builds for one instance.
__global__ void k_plan_fill(int64_t* __restrict__ chunk, int nchunks, int64_t per, int64_t ns,
                            int* __restrict__ inst) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < nchunks) chunk[c] = (int64_t)c * per;
    if (c == nchunks) chunk[c] = ns;
    if (c == 0) {
        inst[0] = 0;
        inst[1] = nchunks;
    }
}

// corr (K, T+1) complex128 from the correlation rows b[t][nv] (slot 3k, 3k+1).
__global__ void k_corr_out(const double* __restrict__ b, int nv, int K, int T,
                           double* __restrict__ corr) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t tot = (int64_t)K * (T + 1);
    if (i >= tot) return;
    const int k = (int)(i / (T + 1)), t = (int)(i % (T + 1));
    corr[i * 2 + 0] = b[(int64_t)t * nv + 3 * k + 0];
    corr[i * 2 + 1] = b[(int64_t)t * nv + 3 * k + 1];
}

// Carves aligned pieces from [base, base + cap).
struct Carve {
    unsigned char* base;
    size_t cap, used = 0;
    void* take(size_t bytes) {
        const size_t off = (used + 255) & ~size_t(255);
        TBG_REQUIRE(off + bytes <= cap, "internal: scratch carve past its end");
        used = off + bytes;
        return base + off;
    }
};

static size_t round256(size_t b) { return (b + 255) & ~size_t(255); }

template <typename R>
static void corr_dos_static(tbg_handle* h, const tbg_ham* H, const double* bes, int nb,
                            const double* psi0, int K, int T, int chk_step, double chk_thr,
                            double* corr, int* fail_step) {
    using C = typename CplxT<R>::T;
    unsigned char* base = static_cast<unsigned char*>(scratch_ptr(h->device));
    TBG_REQUIRE(base != nullptr, "the scratch array of this device was not set up");
    // the chunk plan of make_plan for one instance
    const int64_t ns = H->nslices;
    TBG_REQUIRE(ns >= 1, "empty Hamiltonian");
    const int target = chunk_target(h, ns);
    int64_t per = (ns + target - 1) / target;
    per = std::max<int64_t>(4, (per + 3) / 4 * 4);
    const int nchunks = (int)((ns + per - 1) / per);
    const int Kp = slots_for(K);
    const int64_t n_pad = ns * 64;
    const size_t vbytes = (size_t)n_pad * Kp * sizeof(C);
    const int nv = 3 * Kp;
    const size_t need = 4 * round256(vbytes) +
                        round256((size_t)nchunks * nv * sizeof(double) + 64) +
                        round256((size_t)(nchunks + 1) * sizeof(int64_t)) +
                        round256(2 * sizeof(int)) + round256(sizeof(int) * 4) +
                        round256(((size_t)(T + 1) * nv * sizeof(double) + 15) / 16 * 16);
    if (need > kScratchBytes)
        throw Error{TBG_EINVAL,
                    "sample too large for tbg_corr_dos_noalloc: the run needs " +
                        std::to_string(need) + " bytes of device scratch, the array holds " +
                        std::to_string(kScratchBytes) + " (use tbg_corr_dos)"};
    const hipStream_t st = h->stream;
    TimingHold hold(h);    // no timing events either
    Carve cv{base, kScratchBytes};
    Engine<R> e;
    e.h = h;
    e.M = view_of(H);
    e.Kp = Kp;
    const int V = vectors_per_lane<R>(Kp);
    e.lgG = 0;
    while ((1 << e.lgG) * V < Kp) ++e.lgG;
    e.n_pad = n_pad;
    e.vbytes = vbytes;
    e.X.borrow(cv.take(vbytes), vbytes);
    e.P.borrow(cv.take(vbytes), vbytes);
    e.Y.borrow(cv.take(vbytes), vbytes);
    void* psi = cv.take(vbytes);
    const size_t pbytes = (size_t)nchunks * nv * sizeof(double) + 64;
    e.part.borrow(cv.take(pbytes), pbytes);
    e.plan.nslices = ns;
    e.plan.nchunks = nchunks;
    e.plan.chunk.borrow(cv.take((size_t)(nchunks + 1) * sizeof(int64_t)),
                        (size_t)(nchunks + 1) * sizeof(int64_t));
    e.plan.inst_chunk_dev.borrow(cv.take(2 * sizeof(int)), 2 * sizeof(int));
    int* counter = static_cast<int*>(cv.take(4 * sizeof(int)));
    const size_t cbytes = ((size_t)(T + 1) * nv * sizeof(double) + 15) / 16 * 16;
    double* cbuf = static_cast<double*>(cv.take(cbytes));
    hipLaunchKernelGGL(k_plan_fill, dim3((unsigned)((nchunks + 1 + 255) / 256)), dim3(256), 0,
                       st, e.plan.chunk.template as<int64_t>(), nchunks, per, ns,
                       e.plan.inst_chunk_dev.template as<int>());
    TBG_HIP(hipGetLastError());
    // the array keeps the previous call's contents: clear what is summed
    // into or read before it is written
    launch_fill_zero(e.X.p, vbytes, st);
    launch_fill_zero(e.P.p, vbytes, st);
    launch_fill_zero(e.Y.p, vbytes, st);
    launch_fill_zero(psi, vbytes, st);
    launch_fill_zero(counter, 16, st);
    launch_fill_zero(cbuf, cbytes, st);
    launch_pack<R>(psi0, K, H->n, psi, Kp, 0, H->perm_ptr(), st);
    TBG_HIP(hipMemcpyAsync(e.X.p, psi, vbytes, hipMemcpyDeviceToDevice, st));
    auto reduce = [&](const int* ranges) {
        launch_reduce(e.part.template as<double>(), nv, ranges, 1, cbuf, 0, counter, 0, st);
    };
    // t = 0: <psi0|psi0> (tbpm.f90:58-62)
    launch_dotnorm<R>(e.plan, psi, psi, Kp, Kp, e.part.template as<double>(), st);
    reduce(e.plan.inst_chunk_dev.template as<int>());
    const StepPoly sp = time_poly(bes, nb);
    for (int t = 1; t <= T; ++t) {
        e.step(sp, 0, 0, F_DOT | F_NORM, psi, counter);
        reduce(e.epi_range());
    }
    hipLaunchKernelGGL(k_corr_out, dim3((unsigned)(((int64_t)K * (T + 1) + 255) / 256)),
                       dim3(256), 0, st, cbuf, nv, K, T, corr);
    TBG_HIP(hipGetLastError());
    // norm check of tbpm_dos (tbpm.f90:73-78): norm_ref after step 1, the
    // state of step chk_step - 1 checked at loop index chk_step (2..T)
    if (chk_step >= 2 && chk_step <= T) {
        double nref[64], nchk[64];
        for (int k = 0; k < K; ++k) {
            TBG_HIP(hipMemcpyAsync(&nref[k], cbuf + (size_t)1 * nv + 3 * k + 2, sizeof(double),
                                   hipMemcpyDeviceToHost, st));
            TBG_HIP(hipMemcpyAsync(&nchk[k], cbuf + (size_t)(chk_step - 1) * nv + 3 * k + 2,
                                   sizeof(double), hipMemcpyDeviceToHost, st));
        }
        TBG_HIP(hipStreamSynchronize(st));
        for (int k = 0; k < K; ++k) {
            const double d = std::fabs(nchk[k] - nref[k]);
            if (std::isnan(d) || !std::isfinite(d) || d > chk_thr) {
                if (fail_step) *fail_step = chk_step;
                throw Error{TBG_ENORM, "wave function norm broken at step " +
                                           std::to_string(chk_step) + " (state " +
                                           std::to_string(k) +
                                           "); check the Hermiticity of the Hamiltonian or "
                                           "increase the rescale factor"};
            }
        }
    }
    TBG_HIP(hipStreamSynchronize(st));
}

}  // namespace tbg

using namespace tbg;

int tbg_corr_dos_noalloc(tbg_handle* h, const tbg_ham* H, const double* bes, int n_bes,
                         const double* psi0, int K, int T, int chk_step, double chk_thr,
                         double* corr, int* fail_step) {
    try {
        TBG_REQUIRE(h != nullptr, "null handle");
        TBG_HIP(hipSetDevice(h->device));
        TBG_REQUIRE(H != nullptr && bes != nullptr && psi0 != nullptr && corr != nullptr,
                    "null argument");
        TBG_REQUIRE(H->h == nullptr || H->h->device == h->device,
                    "Hamiltonian on another device");
        TBG_REQUIRE(n_bes >= 3, "need >= 3 Bessel terms");
        TBG_REQUIRE(K >= 1 && K <= 64, "K must be in [1, 64]");
        TBG_REQUIRE(T >= 1, "T must be >= 1");
        if (H->precision == TBG_FP32)
            corr_dos_static<float>(h, H, bes, n_bes, psi0, K, T, chk_step, chk_thr, corr,
                                   fail_step);
        else
            corr_dos_static<double>(h, H, bes, n_bes, psi0, K, T, chk_step, chk_thr, corr,
                                    fail_step);
        set_error("");
        return TBG_OK;
    } catch (const Error& e) {
        set_error(e.msg);
        return e.code;
    } catch (const std::exception& e) {
        set_error(e.what());
        return TBG_EINVAL;
    }
}
