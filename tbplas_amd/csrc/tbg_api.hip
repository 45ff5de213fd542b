// C ABI of the library (declared in include/tbplas_amd.h) and the DOS,
// step, polynomial and batched drivers.
#include <atomic>
#include <cmath>
#include <cstring>
#include <exception>
#include <memory>
#include <new>

#include "tbg_cond.h"
#include "tbg_csr.h"
#include "tbg_engine.h"

namespace tbg {

static thread_local std::string g_err;
void set_error(const std::string& m) { g_err = m; }

template <typename F>
static int guard(F&& f) {
    try {
        f();
        g_err.clear();
        return TBG_OK;
    } catch (const Error& e) {
        set_error(e.msg);
        return e.code;
    } catch (const std::bad_alloc&) {
        set_error("host out of memory");
        return TBG_EOOM;
    } catch (const std::exception& e) {
        set_error(e.what());
        return TBG_EINVAL;
    }
}

static void check_handle(tbg_handle* h) {
    TBG_REQUIRE(h != nullptr, "null handle");
    TBG_HIP(hipSetDevice(h->device));
}

// Norm check of the reference (funcs.f90:175-192): |n - n_ref| > thr or NaN.
static bool norm_broken(double n, double ref, double thr) {
    const double d = std::fabs(n - ref);
    return std::isnan(d) || !std::isfinite(d) || d > thr;
}

}  // namespace tbg

// A DOS correlation run in progress: set up once, advanced in pieces, read
// out at the end (tbg_dos_begin / tbg_run_advance / tbg_dos_end).
struct tbg_run {
    virtual ~tbg_run() = default;
    virtual void advance(int nsteps) = 0;
    virtual void end(int chk_step, double chk_thr, double* const* corr, int* fail_inst,
                     int* fail_step) = 0;
    // the byte counts of the run's tile-pass tables (zeros without them)
    virtual void sched(int64_t* parts, int* row_bytes) const = 0;
    tbg_handle* h = nullptr;
    int T = 0, done = 0;
    // time index of the norm reference of the wave-function check: 1 for
    // DOS (tbpm.f90:73), 0 for LDOS (tbpm.f90:170)
    int norm_ref_t = 1;
};

namespace tbg {

// Three-term tables over a batch's block-diagonal concatenation when every
// instance has spatial tiles: instance i's tiles shifted to its first row,
// blocks never spanning two instances (so the per-tile correlation sums
// stay per instance).  Kept on the handle for the last two batches, since
// the autotuner starts several runs of the same batch.
template <typename R>
static std::shared_ptr<Blocking> batch_blocking(tbg_handle* h, int ni, const tbg_ham* const* hams,
                                                const Concat& cat,
                                                const std::vector<int64_t>& inst_slices, int K) {
    if (!h->blocking || ni < 2) return nullptr;
    for (int i = 0; i < ni; ++i)
        if (hams[i]->tiles.size() < 2) return nullptr;
    const int row_bytes = slots_for(K) * (int)sizeof(typename CplxT<R>::T);
    std::string key = std::to_string(row_bytes) + "|" + std::to_string(h->block_rows) + "|" +
                      std::to_string(h->max_lds) + "|" + std::to_string(h->blocking);
    for (int i = 0; i < ni; ++i) key += "|" + std::to_string(hams[i]->uid);
    for (const auto& kv : h->batch_blk)
        if (kv.first == key) return kv.second;
    std::vector<std::vector<int64_t>> segs((size_t)ni);
    int64_t padded = 0;
    for (int i = 0; i < ni; ++i) {
        const int64_t off = inst_slices[i] * 64;
        for (int64_t t : hams[i]->tiles) segs[i].push_back(off + t);
        padded += hams[i]->padded;
    }
    BlockSrc src;
    const MatView mv = cat.view();
    src.cols = mv.cols;
    src.vals = mv.vals;
    src.sw = mv.sw;
    src.soff = mv.soff;
    src.nslices = cat.nslices;
    src.padded = padded;
    src.nrows = cat.nslices * 64;
    src.vsz = cat.real_vals ? sizeof(R) : sizeof(typename CplxT<R>::T);
    auto B = build_blocking(h, src, segs, row_bytes, h->blocking - 1);
    h->batch_blk.emplace_back(key, B);
    if (h->batch_blk.size() > 2) h->batch_blk.erase(h->batch_blk.begin());
    return B;
}

template <typename R>
struct DosRun : tbg_run {
    int n_inst = 0, K = 0, nv = 0;
    int64_t stride = 0;
    std::vector<int64_t> inst_slices;
    Concat cat;
    Engine<R> e;
    DevBuf psi, cbuf, counter;
    StepPoly sp;

    DosRun(tbg_handle* hh, int ni, const tbg_ham* const* hams, const double* bes, int nb,
           const double* const* psi0, int k, int t) {
        h = hh;
        n_inst = ni;
        K = k;
        T = t;
        inst_slices.assign(1, 0);
        for (int i = 0; i < ni; ++i) inst_slices.push_back(inst_slices.back() + hams[i]->nslices);
        MatView mv;
        if (ni == 1) {
            mv = view_of(hams[0]);
        } else {
            concat_hams(h, ni, hams, cat);  // block-diagonal batch
            mv = cat.view();
        }
        std::shared_ptr<Blocking> pre;
        if (ni > 1) pre = batch_blocking<R>(h, ni, hams, cat, inst_slices, K);
        e.init(h, mv, inst_slices, K, chunk_target(h, mv.nslices), ni == 1 ? hams[0] : nullptr,
               pre);
        const int Kp = e.Kp;
        nv = 3 * Kp;
        psi.alloc(e.vbytes);
        launch_fill_zero(psi.p, e.vbytes, h->stream);
        for (int i = 0; i < ni; ++i) {
            char* base = reinterpret_cast<char*>(psi.p) +
                         (size_t)inst_slices[i] * 64 * Kp * sizeof(typename CplxT<R>::T);
            launch_pack<R>(psi0[i], K, hams[i]->n, base, Kp, 0, hams[i]->perm_ptr(), h->stream);
        }
        TBG_HIP(hipMemcpyAsync(e.X.p, psi.p, e.vbytes, hipMemcpyDeviceToDevice, h->stream));
        stride = (int64_t)(T + 1) * nv;
        cbuf.alloc((size_t)ni * stride * sizeof(double));
        TBG_HIP(hipMemsetAsync(cbuf.p, 0, cbuf.bytes, h->stream));
        counter.alloc(sizeof(int));
        TBG_HIP(hipMemsetAsync(counter.p, 0, sizeof(int), h->stream));
        // t = 0: <psi0|psi0> (tbpm.f90:58-62)
        launch_dotnorm<R>(e.plan, psi.p, psi.p, Kp, Kp, e.part.template as<double>(), h->stream);
        reduce(e.plan.inst_chunk_dev.template as<int>());
        sp = time_poly(bes, nb);
        TBG_HIP(hipStreamSynchronize(h->stream));
    }

    // sums the partial rows of ranges[inst] .. ranges[inst + 1] per instance
    void reduce(const int* ranges) {
        launch_reduce(e.part.template as<double>(), nv, ranges, n_inst, cbuf.as<double>(), stride,
                      counter.as<int>(), 0, h->stream);
    }

    void advance(int nsteps) override {
        TBG_REQUIRE(nsteps >= 0 && done + nsteps <= T, "advancing past the last time step");
        e.run(nsteps, sp, 0, 0, F_DOT | F_NORM, psi.p, counter.as<int>(),
              [&]() { reduce(e.epi_range()); });
        done += nsteps;
    }

    void sched(int64_t* parts, int* row_bytes) const override {
        for (int i = 0; i < 4; ++i) parts[i] = e.blk ? e.blk->sched[i] : 0;
        *row_bytes = e.blk ? e.blk->row_bytes : 0;
    }

    void end(int chk_step, double chk_thr, double* const* corr, int* fail_inst,
             int* fail_step) override {
        std::vector<double> hb((size_t)n_inst * stride);
        TBG_HIP(hipMemcpyAsync(hb.data(), cbuf.p, hb.size() * sizeof(double),
                               hipMemcpyDeviceToHost, h->stream));
        TBG_HIP(hipStreamSynchronize(h->stream));
        int rc = TBG_OK;
        std::string msg;
        for (int i = 0; i < n_inst; ++i) {
            std::vector<double> out((size_t)K * (T + 1) * 2);
            const double* b = hb.data() + (size_t)i * stride;
            for (int k = 0; k < K; ++k)
                for (int t = 0; t <= T; ++t) {
                    out[((size_t)k * (T + 1) + t) * 2 + 0] = b[(size_t)t * nv + 3 * k + 0];
                    out[((size_t)k * (T + 1) + t) * 2 + 1] = b[(size_t)t * nv + 3 * k + 1];
                }
            TBG_HIP(hipMemcpyAsync(corr[i], out.data(), out.size() * sizeof(double),
                                   hipMemcpyHostToDevice, h->stream));
            TBG_HIP(hipStreamSynchronize(h->stream));
            // tbpm.f90:73-78: norm_ref after step 1, checked on the state of
            // step chk_step-1 at loop index chk_step (2..T); LDOS
            // (tbpm.f90:170-176): norm_ref of the initial state, loop index
            // chk_step in 1..T.
            if (rc == TBG_OK && chk_step >= norm_ref_t + 1 && chk_step <= done) {
                for (int k = 0; k < K; ++k) {
                    const double ref = b[(size_t)norm_ref_t * nv + 3 * k + 2];
                    const double nn = b[(size_t)(chk_step - 1) * nv + 3 * k + 2];
                    if (norm_broken(nn, ref, chk_thr)) {
                        rc = TBG_ENORM;
                        if (fail_inst) *fail_inst = i;
                        if (fail_step) *fail_step = chk_step;
                        msg = "wave function norm broken at step " + std::to_string(chk_step) +
                              " (state " + std::to_string(k) + ", instance " +
                              std::to_string(i) + "): |" + std::to_string(nn) + " - " +
                              std::to_string(ref) + "| > threshold; check the Hermiticity of "
                              "the Hamiltonian or increase the rescale factor";
                        break;
                    }
                }
            }
        }
        if (rc != TBG_OK) throw Error{rc, msg};
    }
};

static tbg_run* dos_begin(tbg_handle* h, int n_inst, const tbg_ham* const* hams,
                          const double* bes, int nb, const double* const* psi0, int K, int T) {
    if (hams[0]->precision == TBG_FP32)
        return new DosRun<float>(h, n_inst, hams, bes, nb, psi0, K, T);
    return new DosRun<double>(h, n_inst, hams, bes, nb, psi0, K, T);
}

static void corr_dos_impl(tbg_handle* h, int n_inst, const tbg_ham* const* hams,
                          const double* bes, int nb, const double* const* psi0, int K, int T,
                          int chk_step, double chk_thr, double* const* corr, int* fail_inst,
                          int* fail_step, int norm_ref_t = 1) {
    std::unique_ptr<tbg_run> r(dos_begin(h, n_inst, hams, bes, nb, psi0, K, T));
    r->norm_ref_t = norm_ref_t;
    r->advance(T);
    r->end(chk_step, chk_thr, corr, fail_inst, fail_step);
}

template <typename R>
static void step_impl(tbg_handle* h, const tbg_ham* H, const StepPoly& sp, double* psi, int K,
                      uint64_t bwd) {
    Engine<R> e;
    e.init(h, view_of(H), {0, H->nslices}, K, chunk_target(h, H->nslices), H);
    launch_pack<R>(psi, K, H->n, e.X.p, e.Kp, 0, H->perm_ptr(), h->stream);
    e.step(sp, bwd, 0, 0, nullptr, nullptr);
    launch_unpack<R>(e.X.p, e.Kp, 0, H->n, psi, K, H->perm_ptr(), h->stream);
    TBG_HIP(hipStreamSynchronize(h->stream));
}

// tbpm_psi_t (tbpm.f90:1029-1093): psi0 propagated forward T steps; the
// state after t steps is written to psi_t[i] when time_log[i] == t (a
// sorted list; the pointer stays on the last entry once it is reached,
// as in the reference's loop).  The norm of the state entering step
// chk_step is checked against the initial norm.
template <typename R>
static void psi_t_impl(tbg_handle* h, const tbg_ham* H, const double* bes, int nb,
                       const double* psi0, const int64_t* tlog, int n_log, int T, int chk_step,
                       double chk_thr, double* psi_t, int* fail_step) {
    const hipStream_t st = h->stream;
    Engine<R> e;
    e.init(h, view_of(H), {0, H->nslices}, 1, chunk_target(h, H->nslices), H);
    const int Kp = e.Kp, nv = 3 * Kp;
    const int64_t n = H->n;
    launch_pack<R>(psi0, 1, n, e.X.p, Kp, 0, H->perm_ptr(), st);
    DevBuf nbuf((size_t)nv * sizeof(double));
    auto norm_of_state = [&]() {
        launch_dotnorm<R>(e.plan, e.X.p, e.X.p, Kp, Kp, e.part.template as<double>(), st);
        launch_reduce(e.part.template as<double>(), nv, e.plan.inst_chunk_dev.template as<int>(),
                      1, nbuf.as<double>(), 0, nullptr, 0, st);
        double v[3];
        TBG_HIP(hipMemcpyAsync(v, nbuf.p, sizeof(v), hipMemcpyDeviceToHost, st));
        TBG_HIP(hipStreamSynchronize(st));
        return v[2];
    };
    const double norm_ref = norm_of_state();
    const StepPoly sp = time_poly(bes, nb);
    auto log_frame = [&](int i) {
        launch_unpack<R>(e.X.p, Kp, 0, n, psi_t + (size_t)i * n * 2, 1, H->perm_ptr(), st);
    };
    int ptr = 0;
    int t = 1;   // loop index of the reference: the state holds t-1 steps
    while (t <= T) {
        if (tlog[ptr] == t - 1) {
            log_frame(ptr);
            if (ptr < n_log - 1) ++ptr;
        }
        if (t == chk_step) {
            const double nn = norm_of_state();
            if (norm_broken(nn, norm_ref, chk_thr)) {
                if (fail_step) *fail_step = chk_step;
                throw Error{TBG_ENORM, "wave function norm broken at step " +
                                           std::to_string(chk_step) +
                                           "; check the Hermiticity of the Hamiltonian or "
                                           "increase the rescale factor"};
            }
        }
        // run up to the next loop index that logs or checks
        int next = T + 1;
        if (tlog[ptr] + 1 > t && tlog[ptr] + 1 <= T) next = (int)(tlog[ptr] + 1);
        if (chk_step > t && chk_step < next) next = chk_step;
        e.run(next - t, sp, 0, 0, 0, nullptr, nullptr, [] {});
        t = next;
    }
    if (tlog[ptr] == T) log_frame(ptr);
    TBG_HIP(hipStreamSynchronize(st));
}

static void check_psi_args(const tbg_ham* H, const double* psi0, int K, int T) {
    TBG_REQUIRE(H != nullptr, "null Hamiltonian");
    TBG_REQUIRE(psi0 != nullptr, "null initial states");
    TBG_REQUIRE(K >= 1 && K <= 64, "number of states must be in [1, 64]");
    TBG_REQUIRE(T >= 1, "number of time steps must be >= 1");
}

}  // namespace tbg

using namespace tbg;

extern "C" {

int tbg_version(void) { return 10000; }

const char* tbg_last_error(void) { return g_err.c_str(); }

int tbg_create(tbg_handle** out, int device) {
    return guard([&] {
        TBG_REQUIRE(out != nullptr, "null output pointer");
        int ndev = 0;
        TBG_HIP(hipGetDeviceCount(&ndev));
        TBG_REQUIRE(device >= 0 && device < ndev, "device index out of range");
        std::unique_ptr<tbg_handle> h(new tbg_handle());
        h->device = device;
        TBG_HIP(hipSetDevice(device));
        hipDeviceProp_t prop;
        TBG_HIP(hipGetDeviceProperties(&prop, device));
        h->num_cus = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
        h->max_lds = prop.sharedMemPerBlock > 0 ? (int)prop.sharedMemPerBlock : 65536;
        prepare_term3(h->max_lds);
        static_scratch_init(device);   // the allocation-free entry point's array
        TBG_HIP(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
        *out = h.release();
    });
}

int tbg_destroy(tbg_handle* h) {
    return guard([&] {
        if (!h) return;
        (void)hipSetDevice(h->device);
        if (h->stream) {
            (void)hipStreamSynchronize(h->stream);
            (void)hipStreamDestroy(h->stream);
        }
        for (auto& t : h->timed) {
            (void)hipEventDestroy(t.e0);
            (void)hipEventDestroy(t.e1);
        }
        for (hipEvent_t e : h->ev_pool) (void)hipEventDestroy(e);
        delete h;
    });
}

int tbg_set_graph(tbg_handle* h, int enable) {
    return guard([&] {
        check_handle(h);
        h->use_graph = enable ? 1 : 0;
    });
}

int tbg_set_tuning(tbg_handle* h, int blocks_per_cu, int stage) {
    return guard([&] {
        check_handle(h);
        TBG_REQUIRE(blocks_per_cu >= 1 && blocks_per_cu <= 64, "blocks_per_cu must be in [1, 64]");
        h->blocks_per_cu = blocks_per_cu;
        h->stage = stage ? 1 : 0;
    });
}

int tbg_set_blocking(tbg_handle* h, int enable, int max_rows) {
    return guard([&] {
        check_handle(h);
        TBG_REQUIRE(max_rows >= 0, "max_rows must be >= 0");
        TBG_REQUIRE(enable == 0 || enable == 1 || enable == 3 || enable == 4 || enable == 5,
                    "enable must be 0 (off), 1 or 3 (three terms per pass), 4 or 5");
        h->blocking = enable == 0 ? 0 : enable == 1 ? 3 : enable;
        h->block_rows = max_rows;
    });
}

int tbg_set_real_layout(tbg_handle* h, int enable) {
    return guard([&] {
        check_handle(h);
        h->real_layout = enable ? 1 : 0;
    });
}

int tbg_set_kernel_timing(tbg_handle* h, int enable) {
    return guard([&] {
        check_handle(h);
        h->timing = enable ? 1 : 0;
    });
}

int tbg_kernel_times(tbg_handle* h, int max_kinds, int* kinds, double* total_ms,
                     int64_t* counts, int* n_kinds) {
    return guard([&] {
        check_handle(h);
        TBG_REQUIRE(n_kinds != nullptr && max_kinds >= 0, "null argument");
        TBG_REQUIRE(max_kinds == 0 || (kinds && total_ms && counts), "null argument");
        TBG_HIP(hipSetDevice(h->device));
        TBG_HIP(hipStreamSynchronize(h->stream));
        std::vector<int> ks;
        std::vector<double> ms;
        std::vector<int64_t> cs;
        hipError_t err = hipSuccess;
        for (auto& t : h->timed) {
            float e = 0.0f;
            const hipError_t r = hipEventElapsedTime(&e, t.e0, t.e1);
            if (r != hipSuccess && err == hipSuccess) err = r;
            size_t i = 0;
            while (i < ks.size() && ks[i] != t.kind) ++i;
            if (i == ks.size()) {
                ks.push_back(t.kind);
                ms.push_back(0.0);
                cs.push_back(0);
            }
            ms[i] += e;
            cs[i] += 1;
            h->ev_pool.push_back(t.e0);
            h->ev_pool.push_back(t.e1);
        }
        h->timed.clear();
        TBG_HIP(err);
        *n_kinds = (int)ks.size();
        for (int i = 0; i < (int)ks.size() && i < max_kinds; ++i) {
            kinds[i] = ks[i];
            total_ms[i] = ms[i];
            counts[i] = cs[i];
        }
    });
}

int tbg_ham_dxy_count(tbg_handle* h, int64_t n, int64_t m, const double* orb_eng,
                      const int64_t* hop_i, const int64_t* hop_j, double cutoff, int64_t* indptr,
                      int64_t* nnz) {
    return guard([&] {
        check_handle(h);
        TBG_REQUIRE(orb_eng && indptr && nnz && (m == 0 || (hop_i && hop_j)), "null argument");
        DevBuf tmp;
        *nnz = place_count(h, n, m, orb_eng, 1, cutoff, hop_i, hop_j, indptr, tmp);
    });
}

int tbg_ham_dxy_fill(tbg_handle* h, int64_t n, int64_t m, const double* orb_eng,
                     const int64_t* hop_i, const int64_t* hop_j, const double* hop_v,
                     const double* dr, double cutoff, const int64_t* indptr, int64_t* indices,
                     double* hop, double* dx, double* dy) {
    return guard([&] {
        check_handle(h);
        TBG_REQUIRE(orb_eng && indptr && indices && hop && dx && dy &&
                        (m == 0 || (hop_i && hop_j && hop_v && dr)),
                    "null argument");
        dxy_fill(h, n, m, orb_eng, hop_i, hop_j, hop_v, dr, cutoff, indptr, indices, hop, dx, dy);
    });
}

int tbg_ham_csr_count(tbg_handle* h, int64_t n, int64_t m, const double* orb_eng,
                      const int64_t* hop_i, const int64_t* hop_j, const double* hop_v,
                      int64_t* indptr, int64_t* nnz) {
    return guard([&] {
        check_handle(h);
        TBG_REQUIRE(orb_eng && indptr && nnz && (m == 0 || (hop_i && hop_j && hop_v)),
                    "null argument");
        TBG_REQUIRE(n > 0 && m >= 0, "bad sizes");
        *nnz = canon_build(h, n, m, orb_eng, hop_i, hop_j, hop_v, indptr, nullptr, nullptr,
                           nullptr);
    });
}

int tbg_ham_csr_fill(tbg_handle* h, int64_t n, int64_t m, const double* orb_eng,
                     const int64_t* hop_i, const int64_t* hop_j, const double* hop_v,
                     const int64_t* indptr, int64_t* indices, double* data) {
    return guard([&] {
        check_handle(h);
        TBG_REQUIRE(orb_eng && indptr && indices && data && (m == 0 || (hop_i && hop_j && hop_v)),
                    "null argument");
        TBG_REQUIRE(n > 0 && m >= 0, "bad sizes");
        canon_build(h, n, m, orb_eng, hop_i, hop_j, hop_v, nullptr, indptr, indices, data);
    });
}

int tbg_get_rescale(tbg_handle* h, int64_t n, int64_t m, const double* orb_eng,
                    const int64_t* hop_i, const int64_t* hop_j, const double* hop_v,
                    double* rescale) {
    return guard([&] {
        check_handle(h);
        TBG_REQUIRE(orb_eng && rescale && (m == 0 || (hop_i && hop_j && hop_v)), "null argument");
        TBG_REQUIRE(n > 0 && m >= 0, "bad sizes");
        *rescale = get_rescale(h, n, m, orb_eng, hop_i, hop_j, hop_v);
    });
}

// Host code: the graphene supercell of make_graphene_diamond
// (tbplas/materials/graphene.py:16-31) in the order SuperCell builds it:
// orbitals (ia, ib, io) (super.pyx:383-443), per cell the A->B hop then the
// two B->A hops (check_conj, builder/base.py:76-92; build_hop_pbc,
// super.pyx:860-986), hops crossing an open direction after all periodic
// ones (split_pc_hop / build_hop_gen, super.pyx:643-857), positions and
// wrapped-cell shifts as build_orb_pos / _fast_div (super.pyx:496-612),
// same operation order, no contraction into FMA.  Vacancies drop their
// hops and renumber the remaining orbitals.
int tbg_graphene_hops(int64_t L1, int64_t L2, double t, int pbc1, int pbc2, double a,
                      const int64_t* vacancies, int64_t nvac, double* orb_pos, int64_t* hop_i,
                      int64_t* hop_j, double* hop_v, double* dr, int64_t* n_out,
                      int64_t* m_out) {
#pragma clang fp contract(off)
    return guard([&] {
        TBG_REQUIRE(L1 >= 1 && L2 >= 1 && L1 * L2 < (int64_t(1) << 40), "bad supercell size");
        TBG_REQUIRE(orb_pos && hop_i && hop_j && hop_v && dr && n_out && m_out, "null argument");
        TBG_REQUIRE(nvac == 0 || vacancies != nullptr, "null vacancy list");
        const int64_t n0 = 2 * L1 * L2;
        const double gam = 60.0 / 180 * 3.141592653589793;
        const double lat[3][3] = {{a, 0.0, 0.0},
                                  {a * std::cos(gam), a * std::sin(gam), 0.0},
                                  {0.0, 0.0, 1.0}};
        double sc1[3], sc2[3];
        for (int k = 0; k < 3; ++k) {
            sc1[k] = lat[0][k] * (double)L1;
            sc2[k] = lat[1][k] * (double)L2;
        }
        auto pos = [&](int64_t ia, int64_t ib, int io, double* out) {
            const double f = io ? 1 / 3. : 0.0;
            const double ra = f + (double)ia, rb = f + (double)ib, rc = 0.0;
            for (int k = 0; k < 3; ++k) out[k] = ra * lat[0][k] + rb * lat[1][k] + rc * lat[2][k];
        };
        std::vector<char> vac((size_t)n0, 0);
        for (int64_t v = 0; v < nvac; ++v) {
            TBG_REQUIRE(vacancies[v] >= 0 && vacancies[v] < n0, "vacancy index out of range");
            vac[(size_t)vacancies[v]] = 1;
        }
        std::vector<int64_t> new_id((size_t)n0);
        int64_t n = 0;
        for (int64_t o = 0; o < n0; ++o) {
            new_id[(size_t)o] = vac[(size_t)o] ? -1 : n;
            if (!vac[(size_t)o]) {
                pos(o / (2 * L2), (o / 2) % L2, (int)(o % 2), orb_pos + 3 * n);
                ++n;
            }
        }
        // reduced hops (ra, rb, orb_i, orb_j)
        const int hops[3][4] = {{0, 0, 0, 1}, {1, 0, 1, 0}, {0, 1, 1, 0}};
        const bool per[2] = {pbc1 != 0, pbc2 != 0};
        int64_t m = 0;
        for (int pass = 0; pass < 2; ++pass) {          // periodic hops, then free ones
            for (int64_t ia = 0; ia < L1; ++ia)
                for (int64_t ib = 0; ib < L2; ++ib)
                    for (const auto& hp : hops) {
                        const bool free = (!per[0] && hp[0] != 0) || (!per[1] && hp[1] != 0);
                        if (free != (pass == 1)) continue;
                        const int64_t ja0 = ia + hp[0], jb0 = ib + hp[1];
                        if ((!per[0] && ja0 >= L1) || (!per[1] && jb0 >= L2)) continue;
                        const int64_t na = ja0 / L1, nb = jb0 / L2;
                        const int64_t ja = ja0 - na * L1, jb = jb0 - nb * L2;
                        const int64_t i = (ia * L2 + ib) * 2 + hp[2];
                        const int64_t j = (ja * L2 + jb) * 2 + hp[3];
                        if (vac[(size_t)i] || vac[(size_t)j]) continue;
                        double pi[3], pj[3];
                        pos(ia, ib, hp[2], pi);
                        pos(ja, jb, hp[3], pj);
                        for (int k = 0; k < 3; ++k)
                            dr[3 * m + k] = ((pj[k] - pi[k]) + (double)na * sc1[k]) +
                                            (double)nb * sc2[k];
                        hop_i[m] = new_id[(size_t)i];
                        hop_j[m] = new_id[(size_t)j];
                        hop_v[2 * m] = t;
                        hop_v[2 * m + 1] = 0.0;
                        ++m;
                    }
        }
        *n_out = n;
        *m_out = m;
    });
}

// Host code: same operation order and libm cos/sin as set_mag_field
// (tbplas/cython/sample.pyx:18-72), no contraction into FMA.
int tbg_peierls_host(int64_t m, const int64_t* hop_i, const int64_t* hop_j, double* hop_v,
                     const double* dr, const double* orb_pos, double intensity, int gauge) {
#pragma clang fp contract(off)
    return guard([&] {
        TBG_REQUIRE(m >= 0 && (m == 0 || (hop_i && hop_j && hop_v && dr && orb_pos)),
                    "null argument");
        TBG_REQUIRE(gauge >= 0 && gauge <= 2, "gauge must be 0, 1 or 2");
        const double factor = M_PI / 4135.666734;
        for (int64_t ih = 0; ih < m; ++ih) {
            const int64_t ii = hop_i[ih], jj = hop_j[ih];
            double phase;
            if (gauge == 0) {
                const double sy = orb_pos[jj * 3 + 1] + orb_pos[ii * 3 + 1];
                phase = factor * intensity * dr[ih * 3 + 0] * sy;
            } else if (gauge == 1) {
                const double sx = orb_pos[jj * 3 + 0] + orb_pos[ii * 3 + 0];
                phase = -factor * intensity * dr[ih * 3 + 1] * sx;
            } else {
                const double sx = orb_pos[jj * 3 + 0] + orb_pos[ii * 3 + 0];
                const double sy = orb_pos[jj * 3 + 1] + orb_pos[ii * 3 + 1];
                const double p1 = dr[ih * 3 + 0] * sx;
                const double p2 = dr[ih * 3 + 1] * sy;
                phase = factor * 0.5 * intensity * (p1 - p2);
            }
            const double c = std::cos(phase), s = std::sin(phase);
            const double hr = hop_v[ih * 2 + 0], hi = hop_v[ih * 2 + 1];
            const double t1 = hr * c, t2 = hi * s, t3 = hr * s, t4 = hi * c;
            hop_v[ih * 2 + 0] = t1 - t2;
            hop_v[ih * 2 + 1] = t3 + t4;
        }
    });
}

int tbg_ham_create(tbg_handle* h, int64_t n, const int64_t* indptr, const int64_t* indices,
                   const double* hop, const double* dx, const double* dy, double cur_scale,
                   int precision, tbg_ham** out) {
    return tbg_ham_create_ordered(h, n, indptr, indices, hop, dx, dy, cur_scale, precision,
                                  nullptr, nullptr, 0, out);
}

int tbg_ham_create_ordered(tbg_handle* h, int64_t n, const int64_t* indptr,
                           const int64_t* indices, const double* hop, const double* dx,
                           const double* dy, double cur_scale, int precision,
                           const int64_t* perm, const int64_t* tile_ptr, int64_t ntiles,
                           tbg_ham** out) {
    return guard([&] {
        check_handle(h);
        TBG_REQUIRE(out && indptr && indices && hop, "null argument");
        TBG_REQUIRE(n > 0 && n < (int64_t(1) << 31) - 64, "n must be in (0, 2^31 - 64)");
        TBG_REQUIRE(precision == TBG_FP64 || precision == TBG_FP32, "unknown precision");
        TBG_REQUIRE((dx == nullptr) == (dy == nullptr), "dx and dy must be given together");
        std::unique_ptr<tbg_ham> H(new tbg_ham());
        static std::atomic<uint64_t> next_uid{1};
        H->uid = next_uid++;
        H->h = h;
        H->n = n;
        H->precision = precision;
        if (tile_ptr != nullptr) {
            TBG_REQUIRE(ntiles >= 1, "tiles need ntiles >= 1");
            TBG_REQUIRE(tile_ptr[0] == 0 && tile_ptr[ntiles] == n,
                        "tile_ptr must run from 0 to n");
            for (int64_t t = 0; t < ntiles; ++t)
                TBG_REQUIRE(tile_ptr[t + 1] > tile_ptr[t], "tiles must be non-empty and ordered");
            H->tiles.assign(tile_ptr, tile_ptr + ntiles + 1);
        }
        build_sell(h, H.get(), indptr, indices, hop, dx, dy, cur_scale, perm);
        *out = H.release();
    });
}

int tbg_ham_destroy(tbg_ham* H) {
    return guard([&] {
        if (!H) return;
        (void)hipSetDevice(H->h->device);
        delete H;
    });
}

int tbg_ham_info(const tbg_ham* H, int64_t* n, int64_t* nnz, int64_t* padded, int* max_width) {
    return guard([&] {
        TBG_REQUIRE(H != nullptr, "null Hamiltonian");
        if (n) *n = H->n;
        if (nnz) *nnz = H->nnz;
        if (padded) *padded = H->padded;
        if (max_width) *max_width = H->max_width;
    });
}

int tbg_ham_is_real(const tbg_ham* H, int* real) {
    return guard([&] {
        TBG_REQUIRE(H != nullptr && real != nullptr, "null argument");
        *real = H->real_vals ? 1 : 0;
    });
}

int tbg_ham_blocking_info(tbg_handle* h, const tbg_ham* H, int states, int precision,
                          int* nblocks, double* halo) {
    return guard([&] {
        check_handle(h);
        TBG_REQUIRE(H != nullptr && nblocks != nullptr && halo != nullptr, "null argument");
        TBG_REQUIRE(states >= 1 && states <= 64, "states must be in [1, 64]");
        TBG_REQUIRE(precision == TBG_FP64 || precision == TBG_FP32, "unknown precision");
        const int Kp = slots_for(states);
        const int cb = precision == TBG_FP32 ? (int)sizeof(float2) : (int)sizeof(double2);
        auto b = get_blocking(h, H, Kp * cb);
        *nblocks = b ? b->nblocks : 0;
        *halo = b ? b->halo : 0.0;
    });
}

int tbg_ham_blocking_launch(tbg_handle* h, const tbg_ham* H, int states, int precision,
                            int* max_rows, int* threads) {
    return guard([&] {
        check_handle(h);
        TBG_REQUIRE(H != nullptr && max_rows != nullptr && threads != nullptr,
                    "null argument");
        TBG_REQUIRE(states >= 1 && states <= 64, "states must be in [1, 64]");
        TBG_REQUIRE(precision == TBG_FP64 || precision == TBG_FP32, "unknown precision");
        const int Kp = slots_for(states);
        const int eb = precision == TBG_FP32 ? 4 : 8;
        auto b = get_blocking(h, H, Kp * 2 * eb);
        *max_rows = b ? b->nA_max : 0;
        *threads = b ? term3_launch_threads(eb, H->real_vals != 0, Kp, b->nA_max) : 0;
    });
}

int tbg_ham_blocking_bytes(tbg_handle* h, const tbg_ham* H, int states, int precision,
                           int64_t* parts) {
    return guard([&] {
        check_handle(h);
        TBG_REQUIRE(H != nullptr && parts != nullptr, "null argument");
        TBG_REQUIRE(states >= 1 && states <= 64, "states must be in [1, 64]");
        TBG_REQUIRE(precision == TBG_FP64 || precision == TBG_FP32, "unknown precision");
        const int Kp = slots_for(states);
        const int cb = precision == TBG_FP32 ? (int)sizeof(float2) : (int)sizeof(double2);
        auto b = get_blocking(h, H, Kp * cb);
        for (int i = 0; i < 4; ++i) parts[i] = b ? b->sched[i] : 0;
    });
}

static void check_batch(tbg_handle* h, int n_inst, const tbg_ham* const* hams,
                        const double* bes, int n_bes, const double* const* psi0, int K, int T) {
    check_handle(h);
    TBG_REQUIRE(n_inst >= 1 && hams && psi0, "null argument");
    TBG_REQUIRE(bes && n_bes >= 3, "need >= 3 Bessel terms");
    for (int i = 0; i < n_inst; ++i) {
        check_psi_args(hams[i], psi0[i], K, T);
        TBG_REQUIRE(hams[i]->precision == hams[0]->precision,
                    "all instances must share one precision");
    }
}

int tbg_corr_dos(tbg_handle* h, const tbg_ham* H, const double* bes, int n_bes,
                 const double* psi0, int K, int T, int chk_step, double chk_thr, double* corr,
                 int* fail_step) {
    return guard([&] {
        const tbg_ham* hs[1] = {H};
        const double* ps[1] = {psi0};
        double* cs[1] = {corr};
        check_batch(h, 1, hs, bes, n_bes, ps, K, T);
        TBG_REQUIRE(corr != nullptr, "null output");
        corr_dos_impl(h, 1, hs, bes, n_bes, ps, K, T, chk_step, chk_thr, cs, nullptr, fail_step);
    });
}

int tbg_corr_ldos(tbg_handle* h, const tbg_ham* H, const double* bes, int n_bes,
                  const double* psi0, int K, int T, int chk_step, double chk_thr, double* corr,
                  int* fail_step) {
    return guard([&] {
        const tbg_ham* hs[1] = {H};
        const double* ps[1] = {psi0};
        double* cs[1] = {corr};
        check_batch(h, 1, hs, bes, n_bes, ps, K, T);
        TBG_REQUIRE(corr != nullptr, "null output");
        corr_dos_impl(h, 1, hs, bes, n_bes, ps, K, T, chk_step, chk_thr, cs, nullptr, fail_step,
                      0);
    });
}

int tbg_psi_t(tbg_handle* h, const tbg_ham* H, const double* bes, int n_bes, const double* psi0,
              const int64_t* time_log, int n_log, int T, int chk_step, double chk_thr,
              double* psi_t, int* fail_step) {
    return guard([&] {
        check_handle(h);
        check_psi_args(H, psi0, 1, T);
        TBG_REQUIRE(bes && n_bes >= 3, "need >= 3 Bessel terms");
        TBG_REQUIRE(psi_t != nullptr && time_log != nullptr && n_log >= 1, "null output or log");
        for (int i = 0; i < n_log; ++i) {
            TBG_REQUIRE(time_log[i] >= 0 && time_log[i] <= T, "time_log entry out of [0, T]");
            TBG_REQUIRE(i == 0 || time_log[i] > time_log[i - 1],
                        "time_log must be strictly increasing");
        }
        if (H->precision == TBG_FP32)
            psi_t_impl<float>(h, H, bes, n_bes, psi0, time_log, n_log, T, chk_step, chk_thr,
                              psi_t, fail_step);
        else
            psi_t_impl<double>(h, H, bes, n_bes, psi0, time_log, n_log, T, chk_step, chk_thr,
                               psi_t, fail_step);
    });
}

int tbg_corr_dos_batch(tbg_handle* h, int n_inst, const tbg_ham* const* hams, const double* bes,
                       int n_bes, const double* const* psi0, int K, int T, int chk_step,
                       double chk_thr, double* const* corr, int* fail_inst, int* fail_step) {
    return guard([&] {
        check_batch(h, n_inst, hams, bes, n_bes, psi0, K, T);
        TBG_REQUIRE(corr != nullptr, "null output");
        for (int i = 0; i < n_inst; ++i) TBG_REQUIRE(corr[i] != nullptr, "null output");
        corr_dos_impl(h, n_inst, hams, bes, n_bes, psi0, K, T, chk_step, chk_thr, corr,
                      fail_inst, fail_step);
    });
}

int tbg_dos_begin(tbg_handle* h, int n_inst, const tbg_ham* const* hams, const double* bes,
                  int n_bes, const double* const* psi0, int K, int T, tbg_run** out) {
    return guard([&] {
        check_batch(h, n_inst, hams, bes, n_bes, psi0, K, T);
        TBG_REQUIRE(out != nullptr, "null output pointer");
        *out = dos_begin(h, n_inst, hams, bes, n_bes, psi0, K, T);
    });
}

int tbg_run_blocking_bytes(const tbg_run* r, int64_t* parts, int* row_bytes) {
    return guard([&] {
        TBG_REQUIRE(r != nullptr && parts != nullptr && row_bytes != nullptr, "null argument");
        r->sched(parts, row_bytes);
    });
}

int tbg_run_advance(tbg_run* r, int nsteps, double* elapsed_ms) {
    return guard([&] {
        TBG_REQUIRE(r != nullptr, "null run");
        check_handle(r->h);
        hipEvent_t e0 = nullptr, e1 = nullptr;
        TBG_HIP(hipEventCreate(&e0));
        TBG_HIP(hipEventCreate(&e1));
        struct Ev {
            hipEvent_t a, b;
            ~Ev() { (void)hipEventDestroy(a); (void)hipEventDestroy(b); }
        } ev{e0, e1};
        TBG_HIP(hipEventRecord(e0, r->h->stream));
        r->advance(nsteps);
        TBG_HIP(hipEventRecord(e1, r->h->stream));
        TBG_HIP(hipEventSynchronize(e1));
        float ms = 0.f;
        TBG_HIP(hipEventElapsedTime(&ms, e0, e1));
        if (elapsed_ms) *elapsed_ms = ms;
    });
}

int tbg_dos_end(tbg_run* r, int chk_step, double chk_thr, double* const* corr, int* fail_inst,
                int* fail_step) {
    return guard([&] {
        TBG_REQUIRE(r != nullptr && corr != nullptr, "null argument");
        check_handle(r->h);
        r->end(chk_step, chk_thr, corr, fail_inst, fail_step);
    });
}

int tbg_run_destroy(tbg_run* r) {
    return guard([&] {
        if (!r) return;
        (void)hipSetDevice(r->h->device);
        delete r;
    });
}

int tbg_cheb_step(tbg_handle* h, const tbg_ham* H, const double* bes, int n_bes, double* psi,
                  int K, int fwd) {
    return guard([&] {
        check_handle(h);
        check_psi_args(H, psi, K, 1);
        TBG_REQUIRE(bes && n_bes >= 3, "need >= 3 Bessel terms");
        const StepPoly sp = time_poly(bes, n_bes);
        const uint64_t bwd = fwd ? 0ull : ~0ull;
        if (H->precision == TBG_FP32) step_impl<float>(h, H, sp, psi, K, bwd);
        else step_impl<double>(h, H, sp, psi, K, bwd);
    });
}

int tbg_cheb_poly(tbg_handle* h, const tbg_ham* H, const double* coef, int n_coef, double* psi,
                  int K) {
    return guard([&] {
        check_handle(h);
        check_psi_args(H, psi, K, 1);
        TBG_REQUIRE(coef && n_coef >= 3, "need >= 3 coefficients");
        const StepPoly sp = cheb_poly(coef, n_coef, coef, n_coef);
        if (H->precision == TBG_FP32) step_impl<float>(h, H, sp, psi, K, 0);
        else step_impl<double>(h, H, sp, psi, K, 0);
    });
}

int tbg_corr_ac(tbg_handle* h, const tbg_ham* H, const double* bes, int n_bes,
                const double* coef_F, int n_F, const double* coef_omF, int n_omF,
                const double* psi0, int K, int T, int chk_step, double chk_thr, double* corr,
                int* fail_step) {
    return guard([&] {
        check_handle(h);
        check_psi_args(H, psi0, K, T);
        TBG_REQUIRE(H->has_current, "Hamiltonian was prepared without current operators");
        TBG_REQUIRE(bes && n_bes >= 3 && corr, "need >= 3 Bessel terms and an output");
        TBG_REQUIRE(coef_F && coef_omF && n_F >= 3 && n_omF >= 3, "need >= 3 Fermi coefficients");
        TBG_REQUIRE(3 * K <= 64, "at most 21 states per AC call");
        if (H->precision == TBG_FP32)
            corr_ac_impl<float>(h, H, bes, n_bes, coef_F, n_F, coef_omF, n_omF, psi0, K, T,
                                chk_step, chk_thr, corr, fail_step);
        else
            corr_ac_impl<double>(h, H, bes, n_bes, coef_F, n_F, coef_omF, n_omF, psi0, K, T,
                                 chk_step, chk_thr, corr, fail_step);
    });
}

int tbg_corr_dc(tbg_handle* h, const tbg_ham* H, const double* bes, int n_bes,
                const double* psi0, int K, int T, double t_step, const double* energies,
                const int64_t* en_inds, int n_en, int chk_step, double chk_thr, double* dos_corr,
                double* dc_corr, int* fail_step) {
    return guard([&] {
        check_handle(h);
        check_psi_args(H, psi0, K, T);
        TBG_REQUIRE(H->has_current, "Hamiltonian was prepared without current operators");
        TBG_REQUIRE(bes && n_bes >= 3 && dos_corr && dc_corr, "need >= 3 Bessel terms and outputs");
        TBG_REQUIRE(n_en >= 0 && (n_en == 0 || (energies && en_inds)), "bad energy arguments");
        TBG_REQUIRE(2 * K <= 64, "at most 32 states per DC call");
        if (H->precision == TBG_FP32)
            corr_dc_impl<float>(h, H, bes, n_bes, psi0, K, T, t_step, energies, en_inds, n_en,
                                chk_step, chk_thr, dos_corr, dc_corr, fail_step);
        else
            corr_dc_impl<double>(h, H, bes, n_bes, psi0, K, T, t_step, energies, en_inds, n_en,
                                 chk_step, chk_thr, dos_corr, dc_corr, fail_step);
    });
}

int tbg_corr_dyn_pol(tbg_handle* h, const tbg_ham* H, const double* bes, int n_bes,
                     const double* coef_F, int n_F, const double* coef_omF, int n_omF,
                     const double* psi0, int K, int T, const double* q_points, int n_q,
                     const double* orb_pos, int chk_step, double chk_thr, double* corr,
                     int* fail_step) {
    return guard([&] {
        check_handle(h);
        check_psi_args(H, psi0, K, T);
        TBG_REQUIRE(bes && n_bes >= 3 && corr, "need >= 3 Bessel terms and an output");
        TBG_REQUIRE(coef_F && n_F >= 1 && coef_omF && n_omF >= 1, "need Fermi coefficients");
        TBG_REQUIRE(q_points && n_q >= 1 && orb_pos, "need q-points and orbital positions");
        TBG_REQUIRE(2 * K <= 64, "at most 32 states per call");
        if (H->precision == TBG_FP32)
            dyn_pol_impl<float>(h, H, bes, n_bes, coef_F, n_F, coef_omF, n_omF, psi0, K, T,
                                q_points, n_q, orb_pos, chk_step, chk_thr, corr, fail_step);
        else
            dyn_pol_impl<double>(h, H, bes, n_bes, coef_F, n_F, coef_omF, n_omF, psi0, K, T,
                                 q_points, n_q, orb_pos, chk_step, chk_thr, corr, fail_step);
    });
}

int tbg_supercell_hops(tbg_handle* h, const int32_t* pc_hop_ind, const double* pc_hop_eng,
                       int num_hop_pc, const int32_t* dim, const int32_t* pbc, int num_orb_pc,
                       const int32_t* orb_id_pc, int64_t num_orb_sc, const int64_t* vac_id_sc,
                       int64_t num_vac, const double* sc_lattice, const double* sc_orb_pos,
                       int64_t capacity, int64_t* hop_i, int64_t* hop_j, double* hop_v,
                       double* dr, int64_t* num_hop) {
    return guard([&] {
        check_handle(h);
        TBG_REQUIRE(num_hop_pc >= 0 && num_orb_pc >= 1 && num_orb_sc >= 0 && num_vac >= 0,
                    "bad sizes");
        TBG_REQUIRE(dim && pbc && sc_lattice && num_hop && (num_orb_sc == 0 || (orb_id_pc &&
                    sc_orb_pos)) && (num_hop_pc == 0 || (pc_hop_ind && pc_hop_eng)),
                    "null argument");
        TBG_REQUIRE(dim[0] >= 1 && dim[1] >= 1 && dim[2] >= 1, "supercell dimensions must be >= 1");
        TBG_REQUIRE(num_vac == 0 || vac_id_sc != nullptr, "null vacancy list");
        for (int64_t k = 1; k < num_vac; ++k)
            TBG_REQUIRE(vac_id_sc[k] > vac_id_sc[k - 1], "vacancies must be sorted and unique");
        for (int k = 0; k < num_hop_pc; ++k)
            TBG_REQUIRE(pc_hop_ind[k * 5 + 3] >= 0 && pc_hop_ind[k * 5 + 3] < num_orb_pc &&
                            pc_hop_ind[k * 5 + 4] >= 0 && pc_hop_ind[k * 5 + 4] < num_orb_pc,
                        "hopping orbital index out of range");
        for (int64_t k = 0; k < num_orb_sc; ++k)
            for (int x = 0; x < 3; ++x)
                TBG_REQUIRE(orb_id_pc[k * 4 + x] >= 0 && orb_id_pc[k * 4 + x] < dim[x],
                            "orbital cell index out of range");
        TBG_REQUIRE(num_orb_sc + num_vac == (int64_t)dim[0] * dim[1] * dim[2] * num_orb_pc,
                    "orbital count does not match the supercell and its vacancies");
        TBG_REQUIRE(capacity >= 0 && (capacity == 0 || (hop_i && hop_j && hop_v && dr)),
                    "null output");
        *num_hop = supercell_hops(h, pc_hop_ind, pc_hop_eng, num_hop_pc, dim, pbc, num_orb_pc,
                                  orb_id_pc, num_orb_sc, vac_id_sc, num_vac, sc_lattice,
                                  sc_orb_pos, capacity, hop_i, hop_j, hop_v, dr);
    });
}

int tbg_haydock_coef(tbg_handle* h, const tbg_ham* H, const double* psi0, int K, int depth,
                     double rescale, double* coefa, double* coefb) {
    return guard([&] {
        check_handle(h);
        check_psi_args(H, psi0, K, 1);
        TBG_REQUIRE(H->precision == TBG_FP64, "the Haydock recursion needs a complex128 Hamiltonian");
        TBG_REQUIRE(depth >= 1 && coefa && coefb, "need depth >= 1 and outputs");
        haydock_impl(h, H, psi0, K, depth, rescale, coefa, coefb);
    });
}

int tbg_hall_mu(tbg_handle* h, const tbg_ham* H, const double* psi0, int K, int n_kernel,
                int direction, double* mu) {
    return guard([&] {
        check_handle(h);
        check_psi_args(H, psi0, K, 1);
        TBG_REQUIRE(H->has_current, "Hamiltonian was prepared without current operators");
        TBG_REQUIRE(H->precision == TBG_FP64, "the Hall moments need a complex128 Hamiltonian");
        TBG_REQUIRE(n_kernel >= 1 && mu != nullptr, "need n_kernel >= 1 and an output");
        TBG_REQUIRE(direction == 1 || direction == 2, "direction must be 1 (xx) or 2 (xy)");
        hall_mu_impl(h, H, psi0, K, n_kernel, direction, mu);
    });
}

int tbg_quasi_eigenstates(tbg_handle* h, const tbg_ham* H, const double* bes, int n_bes,
                          const double* psi0, int K, int T, double t_step,
                          const double* energies, int n_en, int chk_step, double chk_thr,
                          double* qe, int* fail_step) {
    return guard([&] {
        check_handle(h);
        check_psi_args(H, psi0, K, T);
        TBG_REQUIRE(bes && n_bes >= 3 && qe, "need >= 3 Bessel terms and an output");
        TBG_REQUIRE(n_en >= 1 && energies, "need at least one energy");
        TBG_REQUIRE(2 * K <= 64, "at most 32 states per call");
        // the energies themselves (tbpm.f90:786) through the DC run's index list
        std::vector<int64_t> idx(n_en);
        for (int i = 0; i < n_en; ++i) idx[i] = i;
        std::vector<double> dos((size_t)K * T * 2);
        DevBuf ddos(dos.size() * sizeof(double));
        if (H->precision == TBG_FP32)
            corr_dc_impl<float>(h, H, bes, n_bes, psi0, K, T, t_step, energies, idx.data(), n_en,
                                chk_step, chk_thr, ddos.as<double>(), nullptr, fail_step, qe);
        else
            corr_dc_impl<double>(h, H, bes, n_bes, psi0, K, T, t_step, energies, idx.data(),
                                 n_en, chk_step, chk_thr, ddos.as<double>(), nullptr, fail_step,
                                 qe);
    });
}

}  // extern "C"
