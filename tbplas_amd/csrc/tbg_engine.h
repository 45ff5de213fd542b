// Step driver shared by the correlation entry points.
#pragma once

#include <algorithm>
#include <type_traits>

#include "tbg_block.h"
#include "tbg_prop.h"

namespace tbg {

// Looks up the device address of the allocation-free entry point's scratch
// array on `device` (current), once per device (tbg_static.hip).
void static_scratch_init(int device);

// Matrix operand of a run: one prepared Hamiltonian or the block-diagonal
// concatenation of several.
struct MatView {
    const int* cols = nullptr;
    const void* vals = nullptr;
    const int64_t* soff = nullptr;
    const int* sw = nullptr;
    int64_t nslices = 0;
    int real_vals = 0;       // vals holds R instead of complex<R>
};

inline MatView view_of(const tbg_ham* H) {
    MatView m;
    m.cols = H->cols.as<int>();
    m.vals = H->vals.p;
    m.soff = H->slice_off.as<int64_t>();
    m.sw = H->slice_w.as<int>();
    m.nslices = H->nslices;
    m.real_vals = H->real_vals ? 1 : 0;
    return m;
}

// Persistent blocks for the term kernels: blocks_per_cu per CU, but no
// chunk smaller than one 4-slice tile.
inline int chunk_target(const tbg_handle* h, int64_t nslices) {
    const int64_t by_cu = (int64_t)h->num_cus * h->blocks_per_cu;
    return (int)std::max<int64_t>(1, std::min<int64_t>(by_cu, (nslices + 3) / 4));
}

// Vector slots: V * G >= the vectors needed with G (lanes per row) a power
// of two that divides the wavefront: the smaller of the next power of two
// and the next 3 * 2^j (three states per lane), e.g. 24 slots for the 24
// vectors of an 8-sample AC run instead of 32.
inline int slots_for(int k_needed) {
    int p2 = 1;
    while (p2 < k_needed) p2 <<= 1;
    int p3 = 3;
    while (p3 < k_needed) p3 <<= 1;
    return p3 < p2 ? p3 : p2;
}

// Polynomial of one step: terms v_0..v_{N-1} with v_1 = alpha1 H v_0 and
// v_m = alpha H v_{m-1} + beta v_{m-2}; output sum_m c[cls][m] v_m.
struct StepPoly {
    std::vector<double> c[2];
    int N = 0;
    double ar1 = 0, ai1 = 0, ar = 0, ai = 0, beta = 0;
};

// exp(-i H~ tau) from the Bessel series (propagation.f90:12-49):
// c_0 = J_0, c_m = 2 J_m, alpha1 = -i, alpha = -2i, beta = 1.
inline StepPoly time_poly(const double* bes, int nb) {
    StepPoly p;
    p.N = nb;
    p.c[0].resize(nb);
    for (int m = 0; m < nb; ++m) p.c[0][m] = m == 0 ? bes[0] : 2.0 * bes[m];
    p.c[1] = p.c[0];
    p.ar1 = 0.0; p.ai1 = -1.0; p.ar = 0.0; p.ai = -2.0; p.beta = 1.0;
    return p;
}

// sum_m c_m T_m(H~) (propagation.f90:52-82): alpha1 = 1, alpha = 2, beta = -1.
inline StepPoly cheb_poly(const double* c0, int n0, const double* c1, int n1) {
    StepPoly p;
    p.N = std::max(n0, n1);
    p.c[0].assign(p.N, 0.0);
    p.c[1].assign(p.N, 0.0);
    for (int m = 0; m < n0; ++m) p.c[0][m] = c0[m];
    for (int m = 0; m < n1; ++m) p.c[1][m] = c1[m];
    p.ar1 = 1.0; p.ai1 = 0.0; p.ar = 2.0; p.ai = 0.0; p.beta = -1.0;
    return p;
}

template <typename R>
struct Engine {
    tbg_handle* h = nullptr;
    MatView M;
    Plan plan;
    int Kp = 1, lgG = 0, lgG3 = 0;   // lanes per row: log2, one-term and three-term kernels
    int64_t n_pad = 0;
    DevBuf X, P, Y, part;
    size_t vbytes = 0;
    // Temporal blocking (three terms per pass) and its two extra recurrence
    // buffers, when the single Hamiltonian of the run has spatial tiles.
    std::shared_ptr<Blocking> blk;
    DevBuf Q, S, tile_range;
    bool epi_tiles = false;

    // H: the run's single Hamiltonian (its spatial tiles enable three terms
    // per pass); pre: tables built elsewhere (a batch's concatenation).
    void init(tbg_handle* hh, const MatView& m, const std::vector<int64_t>& inst_slices,
              int k_needed, int chunk_target, const tbg_ham* H = nullptr,
              std::shared_ptr<Blocking> pre = nullptr) {
        h = hh;
        M = m;
        Kp = slots_for(k_needed);
        TBG_REQUIRE(Kp <= 64, "too many vectors for one launch (at most 64 slots)");
        const int V = vectors_per_lane<R>(Kp);
        lgG = 0;
        while ((1 << lgG) * V < Kp) ++lgG;
        lgG3 = 0;
        while ((1 << lgG3) * term3_vectors_per_lane(Kp) < Kp) ++lgG3;
        make_plan(h, plan, inst_slices, chunk_target);
        n_pad = m.nslices * 64;
        vbytes = (size_t)n_pad * Kp * sizeof(typename CplxT<R>::T);
        X.alloc(vbytes);
        P.alloc(vbytes);
        Y.alloc(vbytes);
        launch_fill_zero(X.p, vbytes, h->stream);
        launch_fill_zero(P.p, vbytes, h->stream);
        launch_fill_zero(Y.p, vbytes, h->stream);
        blk.reset();
        if (pre) blk = pre;
        else if (H != nullptr) blk = get_blocking(h, H, Kp * (int)sizeof(typename CplxT<R>::T));
        int nparts = plan.nchunks;
        if (blk) {
            TBG_REQUIRE(blk->row_bytes == Kp * (int)sizeof(typename CplxT<R>::T) &&
                            (int)blk->seg_block.size() == (int)inst_slices.size(),
                        "internal: blocking tables of another shape");
            Q.alloc(vbytes);
            S.alloc(vbytes);
            launch_fill_zero(Q.p, vbytes, h->stream);
            launch_fill_zero(S.p, vbytes, h->stream);
            // partial sums of a pair pass: one row per tile; instance i owns
            // tiles [seg_block[i], seg_block[i+1])
            tile_range.alloc(blk->seg_block.size() * sizeof(int));
            TBG_HIP(hipMemcpyAsync(tile_range.p, blk->seg_block.data(),
                                   blk->seg_block.size() * sizeof(int), hipMemcpyHostToDevice,
                                   h->stream));
            TBG_HIP(hipStreamSynchronize(h->stream));
            nparts = std::max(nparts, blk->nblocks);
        }
        part.alloc((size_t)nparts * 3 * Kp * sizeof(double) + 64);
        epi_tiles = false;
    }

    // The partial-row range (per instance) that the last step's F_DOT /
    // F_NORM epilogue wrote into part: the chunks of the one-term kernel, or
    // the tiles [0, nblocks) when a pair pass ended the step.
    const int* epi_range() const {
        return epi_tiles ? tile_range.as<int>() : plan.inst_chunk_dev.template as<int>();
    }

    TermArgs base_args(uint64_t bwd, uint64_t cls) const {
        TermArgs a{};
        a.cols = M.cols;
        a.vals = M.vals;
        a.soff = M.soff;
        a.sw = M.sw;
        a.chunk = plan.chunk.as<int64_t>();
        a.nslices = M.nslices;
        a.K = Kp;
        a.lgG = lgG;
        a.bwd_mask = bwd;
        a.cls_mask = cls;
        a.nv = 3 * Kp;
        a.stage = h->stage;
        a.real_vals = M.real_vals;
        return a;
    }

    // T = depth + 1 (3 to 5) terms per pass: terms 1 .. T (PM_M1), then
    // passes of T; the rest r of the N - 1 - T terms after them end the
    // step with the epilogue's sums (per tile): r = 0 as the last full pass,
    // r = 2 .. T - 1 as one pass of r terms, r = 1 with the one-term kernel.
    // The step output receives the terms in the order m = 0, 1, 2, ... as in
    // the one-term schedule, so the states are bit-identical to it; the
    // epilogue's sums of a pass are grouped by tile instead of by chunk.
    // 21 Bessel terms: 5 launches with T = 4, 7 with T = 3 (the last a pair).
    void step_blocked(const StepPoly& sp, uint64_t bwd, uint64_t cls, int epi,
                      const void* psi0, int* counter) {
        const int N = sp.N;   // terms m = 1 .. N-1
        const int T = blk->depth + 1;
        PassArgs base{};
        base.meta = blk->meta.as<BlockMeta>();
        base.rows = blk->rows.as<int>();
        base.colL = blk->colL.as<uint16_t>();
        base.colG = blk->colG.as<int>();
        base.valG = blk->valG.p;
        base.acc = Y.p;
        base.ar[0] = base.ar[1] = sp.ar;
        base.ai[0] = base.ai[1] = sp.ai;
        base.ar1[0] = base.ar1[1] = sp.ar1;
        base.ai1[0] = base.ai1[1] = sp.ai1;
        base.beta = sp.beta;
        base.bwd_mask = bwd;
        base.cls_mask = cls;
        base.K = Kp;
        base.lgG = lgG3;
        base.nA_max = blk->nA_max;
        base.nB_max = blk->nB_max;
        base.nv = 3 * Kp;
        auto pass = [&](int mode, int terms, int m, void* p0, void* p1, void* q1, void* q2) {
            PassArgs t = base;
            t.p0 = p0;
            t.p1 = p1;
            t.q1 = q1;
            t.q2 = q2;
            for (int q = 0; q < 7; ++q) {
                const int j = m - 2 + q;   // c_{m-2} .. c_{m+4}, those in the series
                t.c[q][0] = (j >= 0 && j < N) ? sp.c[0][j] : 0.0;
                t.c[q][1] = (j >= 0 && j < N) ? sp.c[1][j] : 0.0;
            }
            if (mode == PM_M1) t.counter = counter;
            if (mode == PM_NORM || mode == PM_DOTNORM) {
                t.psi0 = psi0;
                t.part = part.as<double>();
            }
            timed_launch(h, kTimePass + 10 * mode + terms, [&] {
                launch_pass<R>(mode, terms, M.real_vals != 0, t, *blk, h->stream);
            });
        };
        // v_{m-2}, v_{m-1} and two free buffers; X (v_0) is free after the
        // first pass and Y receives the output (X and Y swap at the end)
        void* b0 = P.p;
        void* b1 = Q.p;
        void* b2 = S.p;
        void* b3 = X.p;
        pass(PM_M1, T, 1, nullptr, X.p, b0, b1);
        int m = T + 1;
        const int rem = N - m;   // terms m .. N-1, at least one (N >= T + 2)
        const int tail = rem % T == 0 ? T : rem % T;
        for (int i = 0; i < (rem - tail) / T; ++i, m += T) {
            pass(PM_NORMAL, T, m, b0, b1, b2, b3);
            std::swap(b0, b2);
            std::swap(b1, b3);
        }
        if (tail == 1) {
            TermArgs a = base_args(bwd, cls);
            a.acc = Y.p;
            a.pin = b1;
            a.pprev = b0;
            a.pout = b0;
            a.ar[0] = a.ar[1] = sp.ar;
            a.ai[0] = a.ai[1] = sp.ai;
            a.beta = sp.beta;
            a.cc[0] = sp.c[0][m];
            a.cc[1] = sp.c[1][m];
            a.psi0 = psi0;
            a.part = part.as<double>();
            timed_launch(h, kTimeTerm, [&] {
                launch_term<R>(F_PREV | F_ACC | F_ACCRMW | epi, a, plan.nchunks, h->stream);
            });
            epi_tiles = false;
        } else {
            const int mode = (epi & F_DOT) ? PM_DOTNORM : (epi & F_NORM) ? PM_NORM : PM_LAST;
            pass(mode, tail, m, b0, b1, nullptr, nullptr);
            epi_tiles = mode != PM_LAST;
        }
        std::swap(X, Y);
        parity ^= 1;
    }

    // One step: X holds the input, the result is left in X (buffers swap).
    void step(const StepPoly& sp, uint64_t bwd, uint64_t cls, int epi, const void* psi0,
              int* counter) {
        const int N = sp.N;
        TBG_REQUIRE(N >= 3, "at least 3 expansion terms are required");
        if (blk && N >= blk->depth + 3) {
            step_blocked(sp, bwd, cls, epi, psi0, counter);
            return;
        }
        auto buf = [&](int j) { return (j % 2 == 0) ? X.p : P.p; };
        for (int m = 1; m < N; ++m) {
            TermArgs a = base_args(bwd, cls);
            a.acc = Y.p;
            int fl;
            if (m == 1) {
                a.pin = X.p;
                a.pout = P.p;
                a.ar[0] = a.ar[1] = sp.ar1;
                a.ai[0] = a.ai[1] = sp.ai1;
                a.counter = counter;
                fl = F_STORE;
            } else {
                const bool last = (m == N - 1);
                a.pin = buf(m - 1);
                a.pprev = buf(m - 2);
                a.pout = buf(m);
                a.ar[0] = a.ar[1] = sp.ar;
                a.ai[0] = a.ai[1] = sp.ai;
                a.beta = sp.beta;
                fl = F_PREV | (last ? 0 : F_STORE);
                auto setc = [&](double* dst, int idx) {
                    dst[0] = sp.c[0][idx];
                    dst[1] = sp.c[1][idx];
                };
                if (m == 2) {
                    fl |= F_ACC | F_ACCPREV | F_MID;
                    setc(a.cp, 0); setc(a.cm, 1); setc(a.cc, 2);
                } else if (m % 3 == 2) {
                    fl |= F_ACC | F_ACCRMW | F_ACCPREV | F_MID;
                    setc(a.cp, m - 2); setc(a.cm, m - 1); setc(a.cc, m);
                } else if (last) {
                    if (m % 3 == 0) {
                        fl |= F_ACC | F_ACCRMW;
                        setc(a.cc, m);
                    } else {
                        fl |= F_ACC | F_ACCRMW | F_MID;
                        setc(a.cm, m - 1); setc(a.cc, m);
                    }
                }
                if (last) {
                    fl |= epi;
                    a.psi0 = psi0;
                    a.part = part.as<double>();
                }
            }
            timed_launch(h, kTimeTerm, [&] { launch_term<R>(fl, a, plan.nchunks, h->stream); });
        }
        std::swap(X, Y);
        parity ^= 1;
    }

    // Number of X/Y swaps so far, mod 2: a captured graph is only valid at
    // the parity it was captured at.
    int parity = 0;

    // Captured graph of `graph_steps` (even) consecutive steps; replaying it
    // leaves the X/Y roles unchanged.
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    int graph_steps = 0;
    int graph_parity = 0;

    void drop_graph() {
        if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
        if (graph) (void)hipGraphDestroy(graph);
        graph_exec = nullptr;
        graph = nullptr;
        graph_steps = 0;
    }
    ~Engine() { drop_graph(); }

    template <typename F>
    void capture(int per, const StepPoly& sp, uint64_t bwd, uint64_t cls, int epi,
                 const void* psi0, int* counter, F&& after) {
        drop_graph();
        graph_parity = parity;
        TBG_HIP(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
        try {
            for (int s = 0; s < per; ++s) {
                step(sp, bwd, cls, epi, psi0, counter);
                after();
            }
        } catch (...) {
            hipGraph_t dummy = nullptr;
            (void)hipStreamEndCapture(h->stream, &dummy);
            if (dummy) (void)hipGraphDestroy(dummy);
            throw;
        }
        TBG_HIP(hipStreamEndCapture(h->stream, &graph));
        hipError_t e = hipGraphInstantiate(&graph_exec, graph, nullptr, nullptr, 0);
        if (e != hipSuccess) {
            drop_graph();
            TBG_HIP(e);
        }
        graph_steps = per;
    }

    // Enqueue nsteps steps, each followed by `after` (reductions etc.),
    // replaying the captured graph where possible.
    template <typename F>
    void run(int nsteps, const StepPoly& sp, uint64_t bwd, uint64_t cls, int epi,
             const void* psi0, int* counter, F&& after) {
        int done = 0;
        if (h->use_graph && !h->timing) {
            const int per = std::min(16, nsteps & ~1);
            if (graph_steps == 0 && per >= 2) capture(per, sp, bwd, cls, epi, psi0, counter, after);
            if (graph_steps > 0 && parity != graph_parity && nsteps > 0) {
                // one launched step restores the buffer roles of the graph
                step(sp, bwd, cls, epi, psi0, counter);
                after();
                ++done;
            }
            if (graph_steps > 0) {
                const int reps = (nsteps - done) / graph_steps;
                for (int r = 0; r < reps; ++r) TBG_HIP(hipGraphLaunch(graph_exec, h->stream));
                done += reps * graph_steps;
            }
        }
        for (; done < nsteps; ++done) {
            step(sp, bwd, cls, epi, psi0, counter);
            after();
        }
    }
};

}  // namespace tbg
