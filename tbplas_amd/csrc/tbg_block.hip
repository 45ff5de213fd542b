// Temporal blocking of the Chebyshev recurrence on MI355X.
//
// The one-term kernel reads v_{m-1} (gathered) and v_{m-2} and writes v_m
// for every term, ~3.6 vector streams per term.  Here one workgroup owns a
// block-tile R of rows (a run of the caller's spatial tiles) and forms T = 3
// or 4 terms in one pass (T - 1 halo rings E1, E2, E3 of R):
//   phase 1  v_m     on R + E1 .. E_{T-1}  from v_{m-1} (HBM/L2) and v_{m-2}
//   phase 2  v_{m+1} on R + E1 .. E_{T-2}  from v_m (LDS) and v_{m-1}
//   phase 3  v_{m+2} on R (+ E1 if T = 4)  from v_{m+1} (LDS) and v_m (LDS)
//   phase 4  v_{m+3} on R (T = 4)          from v_{m+2} (LDS) and v_{m+1}
// Only the last two terms of the owned rows are written (the next pass
// needs exactly those two), and the step output is updated once per pass.
// Per pass and owned row: reads of v_{m-1} and v_{m-2} on the rings (~1.3
// to 1.5 each for compact tiles), two writes, one output read-modify-write:
// about 2.2 streams per term at T = 3 and 1.7 at T = 4.  The halo rows are
// computed redundantly by neighbouring blocks.
//
// The entries of a block's rows are copied into per-block tables (columns,
// values and, for the rows that gather from LDS, positions in the tile),
// column-major, so each phase reads its entries with one coalesced load
// before its gathers instead of following row -> slice -> entry.
// Workgroups are large (8 or 16 waves) because a tile's LDS leaves room for
// one or two per CU.
//
// Every row sum, recurrence step and output update uses the helpers of
// tbg_vec.h in the same order as the one-term kernel, so the states and the
// step output are bit-identical to it; the FP64 correlation sums of a pass
// that ends the step are grouped by tile instead of by chunk.
#include <algorithm>
#include <cstring>
#include <deque>
#include <type_traits>

#include "tbg_block.h"
#include "tbg_vec.h"

namespace tbg {

// ---------------------------------------------------------------------------
// Tables
// ---------------------------------------------------------------------------
// LDS per workgroup the blocking aims at (two workgroups per CU).
constexpr int kBlockLdsTarget = 72 * 1024;

// Tables of one sliced-ELL matrix (device arrays in src) for spatial tiles
// given per segment (absolute row offsets, first = the segment's start,
// last = its end): whole tiles are grouped up to the row target, never
// across a segment end; rows outside every segment are not computed.
std::shared_ptr<Blocking> build_blocking(tbg_handle* h, const BlockSrc& src,
                                         const std::vector<std::vector<int64_t>>& segs,
                                         int row_bytes, int depth) {
    TBG_REQUIRE(depth == 2 || depth == 3, "internal: halo depth must be 2 or 3");
    const int64_t n = src.nrows, ns = src.nslices;
    const size_t vsz = src.vsz;
    std::vector<int> cols((size_t)src.padded), sw((size_t)ns);
    std::vector<int64_t> soff((size_t)ns + 1);
    std::vector<unsigned char> vals((size_t)src.padded * vsz);
    TBG_HIP(hipMemcpyAsync(cols.data(), src.cols, cols.size() * sizeof(int),
                           hipMemcpyDeviceToHost, h->stream));
    TBG_HIP(hipMemcpyAsync(vals.data(), src.vals, vals.size(), hipMemcpyDeviceToHost,
                           h->stream));
    TBG_HIP(hipMemcpyAsync(sw.data(), src.sw, sw.size() * sizeof(int),
                           hipMemcpyDeviceToHost, h->stream));
    TBG_HIP(hipMemcpyAsync(soff.data(), src.soff, soff.size() * sizeof(int64_t),
                           hipMemcpyDeviceToHost, h->stream));
    TBG_HIP(hipStreamSynchronize(h->stream));
    auto width = [&](int r) { return sw[r >> 6]; };
    auto elem = [&](int r, int j) { return soff[r >> 6] + (int64_t)j * 64 + (r & 63); };
    auto col = [&](int r, int j) { return cols[elem(r, j)]; };

    // owned rows per block-tile: whole tiles, up to the row target
    const int lds = std::min(h->max_lds, kBlockLdsTarget);
    // LDS rows (A + B) per owned row: about 2.6 with two rings, 3 with three
    int64_t target = std::max<int64_t>(64, (int64_t)(lds / ((depth == 2 ? 2.6 : 3.0) * row_bytes)));
    // small samples: at least two workgroups per CU
    int64_t covered = 0;
    for (const auto& t : segs) covered += t.back() - t.front();
    target = std::min<int64_t>(target, std::max<int64_t>(64, covered / (2 * (int64_t)h->num_cus)));
    if (h->block_rows > 0) target = std::min<int64_t>(target, h->block_rows);
    // A block-tile whose A and B areas would not fit the LDS budget is split
    // in two (between its tiles when it has several, else in the middle of
    // its rows) and each half tried again; the budget shrinks and the whole
    // build is redone when the widest A and widest B of different blocks
    // together would not fit.
    struct Span {
        size_t seg, i, j;    // tiles [i, j) of segment seg (j = i + 1: part of one)
        int64_t r0, r1;      // rows [r0, r1)
    };
    std::vector<BlockMeta> meta;
    std::vector<int> seg_block;
    std::vector<int> rows, colG;
    std::vector<uint16_t> colL;
    std::vector<unsigned char> valG;
    std::vector<int> mark((size_t)n, -1), loc((size_t)n, 0);
    int nA_max = 0, nB_max = 0;
    int64_t budget = h->max_lds;
    for (int attempt = 0;; ++attempt) {
        meta.clear();
        seg_block.clear();
        rows.clear();
        colG.clear();
        colL.clear();
        valG.clear();
        std::fill(mark.begin(), mark.end(), -1);
        nA_max = nB_max = 0;
        for (size_t sg = 0; sg < segs.size(); ++sg) {
            const std::vector<int64_t>& t = segs[sg];
            seg_block.push_back((int)meta.size());
            std::deque<Span> todo;
            size_t i = 0;
            while (i + 1 < t.size()) {
                size_t j = i + 1;
                while (j + 1 < t.size() && t[j + 1] - t[i] <= target) ++j;
                todo.push_back(Span{sg, i, j, t[i], t[j]});
                i = j;
            }
            while (!todo.empty()) {
                const Span sp = todo.front();
                todo.pop_front();
                const int b = (int)meta.size();
                const int r0 = (int)sp.r0, r1 = (int)sp.r1;
                const size_t base = rows.size();
                for (int r = r0; r < r1; ++r) {
                    mark[r] = b;
                    loc[r] = (int)(rows.size() - base);
                    rows.push_back(r);
                }
                auto ring = [&](size_t from, size_t to) {
                    for (size_t q = from; q < to; ++q) {
                        const int r = rows[q];
                        for (int jj = 0, w = width(r); jj < w; ++jj) {
                            const int c = col(r, jj);
                            if (mark[c] != b) {
                                mark[c] = b;
                                loc[c] = (int)(rows.size() - base);
                                rows.push_back(c);
                            }
                        }
                    }
                };
                size_t ends[4];
                ends[0] = rows.size();
                for (int k = 1; k <= depth; ++k) {   // E1 .. E_depth
                    ring(k == 1 ? base : ends[k - 2], ends[k - 1]);
                    ends[k] = rows.size();
                }
                const int nA = (int)(ends[depth] - base), nB = (int)(ends[depth - 1] - base);
                // (LDS positions are 16 bits: nA < 65536 whenever A fits the LDS)
                if ((int64_t)(nA + nB) * row_bytes > budget || nA > 65535) {
                    for (size_t q = base; q < rows.size(); ++q) mark[rows[q]] = -1;
                    rows.resize(base);
                    if (r1 - r0 < 32) return nullptr;   // a few rows with too wide a halo
                    Span lo = sp, hi = sp;
                    if (sp.j - sp.i >= 2) {
                        const size_t mid = (sp.i + sp.j) / 2;
                        lo.j = mid;
                        lo.r1 = t[mid];
                        hi.i = mid;
                        hi.r0 = t[mid];
                    } else {
                        lo.r1 = hi.r0 = sp.r0 + (sp.r1 - sp.r0) / 2;
                    }
                    todo.push_front(hi);
                    todo.push_front(lo);
                    continue;
                }
                int wA = 0;
                for (size_t q = base; q < rows.size(); ++q) wA = std::max(wA, width(rows[q]));
                BlockMeta m;
                m.offA = (int64_t)base;
                m.offL = (int64_t)colL.size();
                m.offG = (int64_t)colG.size();
                for (int k = 0; k < 4; ++k) m.n[k] = (int)(ends[std::min(k, depth)] - base);
                m.nB = nB;
                m.nA = nA;
                m.w = wA;
                m.r0 = r0;
                meta.push_back(m);
                colL.resize(colL.size() + (size_t)wA * nB);
                colG.resize(colG.size() + (size_t)wA * nA);
                valG.resize(valG.size() + (size_t)wA * nA * vsz, 0);
                uint16_t* cl = colL.data() + m.offL;
                int* cg = colG.data() + m.offG;
                unsigned char* vg = valG.data() + (size_t)m.offG * vsz;
                for (int l = 0; l < nA; ++l) {
                    const int r = rows[base + l];
                    const int w = width(r);
                    for (int jj = 0; jj < wA; ++jj) {
                        const size_t q = (size_t)jj * nA + l;
                        if (jj < w) {
                            cg[q] = col(r, jj);
                            std::memcpy(vg + q * vsz, vals.data() + (size_t)elem(r, jj) * vsz,
                                        vsz);
                        } else {
                            cg[q] = r;       // zero value on the row's own column
                        }
                        if (l < nB)
                            cl[(size_t)jj * nB + l] = (uint16_t)(jj < w ? loc[col(r, jj)] : l);
                    }
                }
                nA_max = std::max(nA_max, nA);
                nB_max = std::max(nB_max, nB);
            }
        }
        seg_block.push_back((int)meta.size());
        if ((int64_t)(nA_max + nB_max) * row_bytes <= h->max_lds) break;
        if (attempt == 3) return nullptr;
        budget = budget * 7 / 8;
    }
    const int nb = (int)meta.size();
    // every LDS index in range: a row of R + E1 + .. + Ek reads positions
    // in R + E1 + .. + E(k+1), which a pass has computed one phase earlier
    for (int b = 0; b < nb; ++b) {
        const BlockMeta& m = meta[b];
        for (int l = 0; l < m.nB; ++l) {
            int k = 0;
            while (l >= m.n[k]) ++k;
            const int lim = m.n[k + 1];
            for (int j = 0; j < m.w; ++j) {
                const int v = colL[m.offL + (int64_t)j * m.nB + l];
                TBG_REQUIRE(v >= 0 && v < lim, "internal: blocking table out of range");
            }
        }
        for (int l = 0; l < m.n[0]; ++l)
            TBG_REQUIRE(rows[m.offA + l] == m.r0 + l, "internal: blocking table out of order");
        for (int64_t q = 0; q < (int64_t)m.w * m.nA; ++q) {
            const int c = colG[m.offG + q];
            TBG_REQUIRE(c >= 0 && c < n, "internal: blocking table out of range");
        }
    }
    auto B = std::make_shared<Blocking>();
    B->row_bytes = row_bytes;
    B->depth = depth;
    B->nblocks = nb;
    B->seg_block = seg_block;
    B->nA_max = nA_max;
    B->nB_max = nB_max;
    B->halo = (double)rows.size() / (double)std::max<int64_t>(covered, 1);
    B->meta.alloc(meta.size() * sizeof(BlockMeta));
    B->rows.alloc(rows.size() * sizeof(int));
    B->colL.alloc(std::max<size_t>(colL.size(), 1) * sizeof(uint16_t));
    B->colG.alloc(std::max<size_t>(colG.size(), 1) * sizeof(int));
    B->valG.alloc(std::max<size_t>(valG.size(), 16));
    TBG_HIP(hipMemcpyAsync(B->meta.p, meta.data(), meta.size() * sizeof(BlockMeta),
                           hipMemcpyHostToDevice, h->stream));
    TBG_HIP(hipMemcpyAsync(B->rows.p, rows.data(), rows.size() * sizeof(int),
                           hipMemcpyHostToDevice, h->stream));
    if (!colL.empty())
        TBG_HIP(hipMemcpyAsync(B->colL.p, colL.data(), colL.size() * sizeof(uint16_t),
                               hipMemcpyHostToDevice, h->stream));
    if (!colG.empty()) {
        TBG_HIP(hipMemcpyAsync(B->colG.p, colG.data(), colG.size() * sizeof(int),
                               hipMemcpyHostToDevice, h->stream));
        TBG_HIP(hipMemcpyAsync(B->valG.p, valG.data(), valG.size(), hipMemcpyHostToDevice,
                               h->stream));
    }
    TBG_HIP(hipStreamSynchronize(h->stream));
    return B;
}

std::shared_ptr<Blocking> get_blocking(tbg_handle* h, const tbg_ham* H, int row_bytes) {
    if (!h->blocking || H->tiles.size() < 2) return nullptr;
    const int depth = h->blocking - 1;
    const bool same = H->blk_row_bytes == row_bytes && H->blk_rows_cap == h->block_rows &&
                      H->blk_depth == depth;
    if (H->blk && same) return H->blk;
    if (H->blk_failed && same) return nullptr;
    H->blk.reset();
    H->blk_failed = false;
    H->blk_row_bytes = row_bytes;
    H->blk_rows_cap = h->block_rows;
    H->blk_depth = depth;
    BlockSrc src;
    src.cols = H->cols.as<int>();
    src.vals = H->vals.p;
    src.sw = H->slice_w.as<int>();
    src.soff = H->slice_off.as<int64_t>();
    src.nslices = H->nslices;
    src.padded = H->padded;
    src.nrows = H->n;
    src.vsz = H->vals.bytes / (size_t)std::max<int64_t>(H->padded, 1);
    auto B = build_blocking(h, src, {H->tiles}, row_bytes, depth);
    if (!B) H->blk_failed = true;
    H->blk = B;
    return B;
}

// ---------------------------------------------------------------------------
// Kernel
// ---------------------------------------------------------------------------
// One lane's rows l = l0, l0 + step, ... < n of a phase when every row has
// NJ <= 4 entries: the row number and the NJ table entries (index idx(j, l),
// value at vals[j * nA + l]) of the next row are loaded before the current
// row's work, so their latency hides behind it.  body(l, t) does the row's
// work from its prefetched entries t (sums via tab_sum, in j order as
// row_sum).  With GR the row numbers come from rows[l], else they are the
// owned run r0 + l (the phases after the first only write owned rows).
template <typename R, bool RV, int NJ>
struct RowTab {
    int gr;
    int col[NJ];
    HVal<R, RV> h[NJ];
};

template <typename R, bool RV, int NJ, bool GR, typename IDX, typename BODY>
__device__ __forceinline__ void rows_prefetched(int l0, int n, int step, const int* rows,
                                                int r0, const unsigned char* vals, int nA,
                                                IDX idx, BODY body) {
    auto load = [&](int l) {
        RowTab<R, RV, NJ> t;
        if constexpr (GR) t.gr = rows[l];
        else t.gr = r0 + l;
#pragma unroll
        for (int u = 0; u < NJ; ++u) {
            t.col[u] = idx(u, l);
            t.h[u].load(vals, u * nA + l);
        }
        return t;
    };
    int l = l0;
    if (l >= n) return;
    RowTab<R, RV, NJ> cur = load(l);
    while (true) {
        const int ln = l + step;
        const bool more = ln < n;
        RowTab<R, RV, NJ> nxt;
        if (more) nxt = load(ln);
        body(l, cur);
        if (!more) break;
        cur = nxt;
        l = ln;
    }
}

// s = sum_j h_j x[col_j] over a prefetched row (gathers first, then the
// products in j order).
template <typename R, int V, bool RV, int NJ, typename GATHER>
__device__ __forceinline__ Vec<R, V> tab_sum(const RowTab<R, RV, NJ>& t, GATHER gather) {
    Vec<R, V> x[NJ];
#pragma unroll
    for (int u = 0; u < NJ; ++u) x[u] = gather(t.col[u]);
    Vec<R, V> s = vzero<R, V>();
#pragma unroll
    for (int u = 0; u < NJ; ++u) mac<R, V, RV>(s, t.h[u], x[u]);
    return s;
}

// NT threads per workgroup: the tile's LDS limits a CU to about two
// workgroups, so they are large (16 waves, or 8 where term3_launch_threads
// says so) to keep enough gathers in flight.  G = K/V lanes share a row,
// NT/G rows per pass of a phase.
//
// T terms m .. m+T-1 per pass (2 <= T <= the tables' depth + 1), phase by
// phase, n[k] = rows of R and the first k halo rings:
//   1: v_m on n[T-1] rows from the gathered v_{m-1} (HBM/L2) and v_{m-2},
//      into LDS area A; the n[T-2] rows of phase 2 also park their v_{m-1}
//      in (the still unused) area B: read next to this tile's gathers of the
//      same rows, it mostly hits L2;
//   2: v_{m+1} on n[T-2] rows from A and the parked v_{m-1}, into B;
//   3: v_{m+2} on n[T-3] rows from B and v_m.  For T = 4 it goes over v_m in
//      A (each row reads its own v_m first, and the phase gathers only from
//      B) and the owned rows' output so far is stored in acc;
//   4: v_{m+3} on the owned rows from A (v_{m+2}) and B (v_{m+1}).
// The last phase works on the owned rows only: acc gets the pass's terms
// (read first, formed from v_0 for PM_M1, or the partial of phase 3), the
// last two terms go to q1, q2 unless the pass ends the step, and the
// ending modes sum the FP64 norm / dot per tile into part.
template <typename R, int V, bool RV, int MODE, int NT, int T>
__global__ void __launch_bounds__(NT) k_block_pass(PassArgs a) {
    static_assert(T >= 2 && T <= 4, "two to four terms per pass");
    static_assert(T > 2 || MODE >= PM_LAST, "a pair only ends a step");
    using C = typename CplxT<R>::T;
    using VT = Vec<R, V>;
    constexpr bool END = MODE >= PM_LAST;
    constexpr bool NORM = MODE == PM_NORM || MODE == PM_DOTNORM;
    constexpr bool DOT = MODE == PM_DOTNORM;
    constexpr bool M1 = MODE == PM_M1;
    extern __shared__ __align__(16) unsigned char tbg_smem3[];
    const int b = remap_chunk(blockIdx.x, gridDim.x);
    if constexpr (M1) {
        // the step counter the last term's epilogue and the reductions read
        if (a.counter != nullptr && blockIdx.x == 0 && threadIdx.x == 0) *a.counter += 1;
    }
    const BlockMeta m = a.meta[b];
    const int* rows = a.rows + m.offA;
    const uint16_t* colL = a.colL + m.offL;
    const int* colG = a.colG + m.offG;
    const int K = a.K;
    unsigned char* ldsA = tbg_smem3;
    unsigned char* ldsB = tbg_smem3 + (size_t)a.nA_max * K * sizeof(C);
    const int lgG = a.lgG;
    const int g = threadIdx.x & ((1 << lgG) - 1);
    const int rsub = threadIdx.x >> lgG;
    const int rpp = NT >> lgG;
    const int k0 = g * V;

    R ar[V], ai[V], cf[6][V];
    R ar1[M1 ? V : 1], ai1[M1 ? V : 1];
#pragma unroll
    for (int v = 0; v < V; ++v) {
        const int k = k0 + v;
        const int cl = (int)((a.cls_mask >> k) & 1ull);
        const bool bwd = ((a.bwd_mask >> k) & 1ull) != 0;
        ar[v] = (R)a.ar[cl];
        ai[v] = bwd ? (R)(-a.ai[cl]) : (R)a.ai[cl];
        if constexpr (M1) {
            ar1[v] = (R)a.ar1[cl];
            ai1[v] = bwd ? (R)(-a.ai1[cl]) : (R)a.ai1[cl];
        }
#pragma unroll
        for (int q = 0; q < 6; ++q) cf[q][v] = (R)a.c[q][cl];
    }
    const R beta = (R)a.beta;
    constexpr size_t vsz = RV ? sizeof(R) : sizeof(C);
    const unsigned char* vals = reinterpret_cast<const unsigned char*>(a.valG) + m.offG * vsz;
    const int w = m.w, nA = m.nA, nB = m.nB, r0 = m.r0;
    const int nR = m.n[0];
    auto lds = [&](unsigned char* area, int l) { return vld<R, V>(area, (int64_t)l * K + k0); };
    auto lds_st = [&](unsigned char* area, int l, const VT& x) {
        vst<R, V>(area, (int64_t)l * K + k0, x);
    };
    // the output before this pass's terms
    auto acc0 = [&](int gr) {
        const int64_t vidx = (int64_t)gr * K + k0;
        VT o;
        if constexpr (M1) {
            o = vzero<R, V>();
            axpy<R, V>(o, cf[1], vld<R, V>(a.p1, vidx));
        } else {
            o = vld<R, V>(a.acc, vidx);
        }
        return o;
    };
    auto none = [&](int, int) { return vzero<R, V>(); };

    // One phase over rows l < n (row numbers from the table with GR, else
    // owned), sums of the row's entries over gather(position); pre(l, gr)
    // runs (and its loads issue) before the gathers, body(l, gr, s, pre).
    // With two or more states per lane, rows of up to four entries (lattice
    // models) prefetch the next row's table entries before the current
    // row's work, so each row waits one memory round trip (+4.5% on
    // 8.4M-orbital graphene, K=8); with one state per lane the workgroup has
    // 16 waves to hide latency and the prefetch's registers cost more than
    // they gain (-2%).
    auto phase = [&](auto grc, int n, auto idx, auto gather, auto pre, auto body) {
        constexpr bool GR = decltype(grc)::value;
        if constexpr (V >= 2) {
            auto piped = [&](auto nj) {
                constexpr int NJ = decltype(nj)::value;
                rows_prefetched<R, RV, NJ, GR>(rsub, n, rpp, rows, r0, vals, nA, idx,
                                               [&](int l, const RowTab<R, RV, NJ>& t) {
                    const VT p = pre(l, t.gr);
                    body(l, t.gr, tab_sum<R, V, RV, NJ>(t, gather), p);
                });
            };
            switch (w) {
                case 1: piped(std::integral_constant<int, 1>{}); return;
                case 2: piped(std::integral_constant<int, 2>{}); return;
                case 3: piped(std::integral_constant<int, 3>{}); return;
                case 4: piped(std::integral_constant<int, 4>{}); return;
                default: break;
            }
        }
        for (int l = rsub; l < n; l += rpp) {
            int gr;
            if constexpr (GR) gr = rows[l];
            else gr = r0 + l;
            const VT p = pre(l, gr);
            body(l, gr, row_sum<R, V, RV>(
                            w, vals, [&](int j) { return idx(j, l); },
                            [&](int j) { return j * nA + l; }, gather), p);
        }
    };
    auto idxG = [&](int j, int l) { return colG[j * nA + l]; };
    auto idxL = [&](int j, int l) { return (int)colL[j * nB + l]; };
    auto gatherG = [&](int c) { return vld<R, V>(a.p1, (int64_t)c * K + k0); };
    auto gatherA = [&](int c) { return lds(ldsA, c); };
    auto gatherB = [&](int c) { return lds(ldsB, c); };
    const std::true_type table_rows{};
    const std::false_type owned_rows{};

    double dre[V], dim_[V], nrm[V];
#pragma unroll
    for (int v = 0; v < V; ++v) { dre[v] = 0.0; dim_[v] = 0.0; nrm[v] = 0.0; }
    // the owned row's new output: stored, and summed by the ending modes
    auto finish = [&](int64_t vidx, const VT& o) {
        vst<R, V>(a.acc, vidx, o);
        if constexpr (DOT) {
            const VT p0 = stream_ld<R, V>(a.psi0, vidx);
#pragma unroll
            for (int v = 0; v < V; ++v) {
                const double pr = p0.re[v], pi = p0.im[v];
                const double orr = o.re[v], oi = o.im[v];
                dre[v] = fma(pr, orr, fma(pi, oi, dre[v]));
                dim_[v] = fma(pr, oi, fma(-pi, orr, dim_[v]));
            }
        }
        if constexpr (NORM) {
#pragma unroll
            for (int v = 0; v < V; ++v) {
                const double orr = o.re[v], oi = o.im[v];
                nrm[v] = fma(orr, orr, fma(oi, oi, nrm[v]));
            }
        }
    };

    // phase 1: v_m
    phase(table_rows, m.n[T - 1], idxG, gatherG,
          [&](int l, int gr) {
              if (l < m.n[T - 2]) lds_st(ldsB, l, vld<R, V>(a.p1, (int64_t)gr * K + k0));
              return vzero<R, V>();
          },
          [&](int l, int gr, const VT& s, const VT&) {
              VT v;
              if constexpr (M1) {
                  v = recur<R, V, false>(ar1, ai1, beta, s, s);
              } else {
                  const VT prev = vld<R, V>(a.p0, (int64_t)gr * K + k0);
                  v = recur<R, V, true>(ar, ai, beta, s, prev);
              }
              lds_st(ldsA, l, v);
          });
    __syncthreads();
    if constexpr (T == 2) {
        // phase 2 (last): v_{m+1} on the owned rows
        phase(owned_rows, nR, idxL, gatherA, [&](int, int gr) { return acc0(gr); },
              [&](int l, int gr, const VT& s, const VT& o0) {
                  const VT v1 = recur<R, V, true>(ar, ai, beta, s, lds(ldsB, l));
                  VT o = o0;
                  axpy<R, V>(o, cf[2], lds(ldsA, l));
                  axpy<R, V>(o, cf[3], v1);
                  finish((int64_t)gr * K + k0, o);
              });
    } else {
        // phase 2: v_{m+1} on n[T-2] rows (each lane reads its own row's
        // parked v_{m-1} before overwriting it)
        phase(owned_rows, m.n[T - 2], idxL, gatherA, none,
              [&](int l, int gr, const VT& s, const VT&) {
                  const VT v1 = recur<R, V, true>(ar, ai, beta, s, lds(ldsB, l));
                  lds_st(ldsB, l, v1);
                  if constexpr (T == 3 && !END) {
                      if (l < nR) vst<R, V>(a.q1, (int64_t)gr * K + k0, v1);
                  }
              });
        __syncthreads();
        if constexpr (T == 3) {
            // phase 3 (last): v_{m+2} on the owned rows
            phase(owned_rows, nR, idxL, gatherB, [&](int, int gr) { return acc0(gr); },
                  [&](int l, int gr, const VT& s, const VT& o0) {
                      const int64_t vidx = (int64_t)gr * K + k0;
                      const VT vm = lds(ldsA, l);
                      const VT v2 = recur<R, V, true>(ar, ai, beta, s, vm);
                      if constexpr (!END) vst<R, V>(a.q2, vidx, v2);
                      VT o = o0;
                      axpy<R, V>(o, cf[2], vm);
                      axpy<R, V>(o, cf[3], lds(ldsB, l));
                      axpy<R, V>(o, cf[4], v2);
                      finish(vidx, o);
                  });
        } else {
            // The owned rows' output after three terms waits for phase 4 in
            // registers: each lane meets its owned rows in the same order in
            // both phases, so its first HELD partials go through a queue
            // with fixed register indices (part[HELD - 1] leaves first),
            // any further ones through acc.  One state per lane: 1024 lanes
            // per tile, about one owned row each, and a longer queue would
            // cost the second workgroup per CU (more than 64 registers).
            constexpr int HELD = V == 1 ? 1 : (V * sizeof(R) <= 8 ? 4 : 2);
            VT part[HELD];
#pragma unroll
            for (int q = 0; q < HELD; ++q) part[q] = vzero<R, V>();
            auto shift = [&]() {
#pragma unroll
                for (int q = HELD - 1; q > 0; --q) part[q] = part[q - 1];
            };
            int held = 0;
            // phase 3: v_{m+2} on R + E1, over v_m in A
            phase(owned_rows, m.n[1], idxL, gatherB,
                  [&](int l, int gr) { return l < nR ? acc0(gr) : vzero<R, V>(); },
                  [&](int l, int gr, const VT& s, const VT& o0) {
                      const VT vm = lds(ldsA, l);
                      const VT v2 = recur<R, V, true>(ar, ai, beta, s, vm);
                      lds_st(ldsA, l, v2);
                      if (l < nR) {
                          const int64_t vidx = (int64_t)gr * K + k0;
                          if constexpr (!END) vst<R, V>(a.q1, vidx, v2);
                          VT o = o0;
                          axpy<R, V>(o, cf[2], vm);
                          axpy<R, V>(o, cf[3], lds(ldsB, l));
                          axpy<R, V>(o, cf[4], v2);
                          if (held < HELD) {
                              shift();
                              part[0] = o;
                              ++held;
                          } else {
                              vst<R, V>(a.acc, vidx, o);
                          }
                      }
                  });
            for (int q = held; q < HELD; ++q) shift();   // first partial to part[HELD - 1]
            __syncthreads();
            // phase 4 (last): v_{m+3} on the owned rows, the output from
            // the queue (or read back from acc, stored by this lane)
            int used = 0;
            phase(owned_rows, nR, idxL, gatherA,
                  [&](int, int gr) {
                      if (used < held) {
                          const VT o = part[HELD - 1];
                          shift();
                          ++used;
                          return o;
                      }
                      return vld<R, V>(a.acc, (int64_t)gr * K + k0);
                  },
                  [&](int l, int gr, const VT& s, const VT& o0) {
                      const int64_t vidx = (int64_t)gr * K + k0;
                      const VT v3 = recur<R, V, true>(ar, ai, beta, s, lds(ldsB, l));
                      if constexpr (!END) vst<R, V>(a.q2, vidx, v3);
                      VT o = o0;
                      axpy<R, V>(o, cf[5], v3);
                      finish(vidx, o);
                  });
        }
    }

    if constexpr (NORM) {
        // Lanes of a wave with equal g hold the same states: butterfly over
        // them, then the waves' sums in wave order (LDS reused: every lane
        // is past its last LDS read at the barrier).
        const int lane = threadIdx.x & 63;
        for (int o = 1 << lgG; o < 64; o <<= 1) {
#pragma unroll
            for (int v = 0; v < V; ++v) {
                dre[v] += __shfl_xor(dre[v], o);
                dim_[v] += __shfl_xor(dim_[v], o);
                nrm[v] += __shfl_xor(nrm[v], o);
            }
        }
        constexpr int NW = NT / 64;
        double* red = reinterpret_cast<double*>(tbg_smem3);   // [NW][3K]
        __syncthreads();
        if ((lane >> lgG) == 0) {
            const int wv = threadIdx.x >> 6;
#pragma unroll
            for (int v = 0; v < V; ++v) {
                const int k = k0 + v;
                red[wv * 3 * K + k * 3 + 0] = dre[v];
                red[wv * 3 * K + k * 3 + 1] = dim_[v];
                red[wv * 3 * K + k * 3 + 2] = nrm[v];
            }
        }
        __syncthreads();
        for (int q = threadIdx.x; q < 3 * K; q += NT) {
            double t = red[q];
            for (int wv = 1; wv < NW; ++wv) t += red[wv * 3 * K + q];
            a.part[(int64_t)b * a.nv + q] = t;
        }
    }
}

// Threads per workgroup for K state slots (V states per lane, G = K/V
// lanes per row) over tiles of at most nA rows: one state per lane always
// takes 16 waves (its registers allow it, and its tiles are the largest),
// three states per lane 8 (their registers would not fit 16 waves).  With
// two states per lane, 16 waves when the widest tile gives every lane two
// rows (nA * G >= 2048), else 8.  Measured on 8.4M-orbital graphene, 3.5 nm
// tiles, step·vectors/s with 8 / 16 waves (nA * G): complex64 K=2 989 / 999
// (2088), K=4 1148 / 1142 (2118), K=8 1278 / 1327 (2492), K=16 959 / 1352
// (4984); complex128 K=2 564 / 501 (1059), K=4 610 / 589 (1246), K=8
// 632 / 643 (2492).  complex128 states with complex values need more than
// the 128 registers a lane has at 16 waves, so that pair stays at 8.
int term3_launch_threads(int elem_bytes, bool real_vals, int K, int nA) {
    const int V = term3_vectors_per_lane(K);
    if (V == 1) return 1024;
    if (V != 2 || (elem_bytes == 8 && !real_vals)) return 512;
    return (int64_t)nA * (K / V) >= 2048 ? 1024 : 512;
}

template <typename R, int V, bool RV, int MODE, int T>
static void launch_f(const PassArgs& a, const Blocking& b, hipStream_t s) {
    // the instances term3_launch_threads can ask for; complex128 states with
    // the dot epilogue after three or four terms need more than the 128
    // registers a lane has at 16 waves, so they take 8
    constexpr bool wide =
        V == 1 || (V == 2 && (sizeof(R) == 4 || (RV && !(MODE == PM_DOTNORM && T >= 3))));
    int nt = term3_launch_threads((int)sizeof(R), RV, a.K, b.nA_max);
    if (!wide && V >= 2) nt = 512;
    size_t lds = (size_t)(b.nA_max + b.nB_max) * a.K * sizeof(typename CplxT<R>::T);
    if constexpr (MODE == PM_NORM || MODE == PM_DOTNORM)   // the waves' partial sums
        lds = std::max(lds, (size_t)(nt / 64) * 3 * a.K * sizeof(double));
    if constexpr (wide) {
        if (nt == 1024) {
            hipLaunchKernelGGL((k_block_pass<R, V, RV, MODE, 1024, T>), dim3(b.nblocks),
                               dim3(1024), lds, s, a);
            TBG_HIP(hipGetLastError());
            return;
        }
    }
    if constexpr (V >= 2) {
        if (nt == 512) {
            hipLaunchKernelGGL((k_block_pass<R, V, RV, MODE, 512, T>), dim3(b.nblocks),
                               dim3(512), lds, s, a);
            TBG_HIP(hipGetLastError());
            return;
        }
    }
    throw Error{TBG_EINVAL, "internal: no tile-pass kernel for this workgroup size"};
}

// Allow every instance of the tile-pass kernel the LDS of one workgroup
// (tables are sized up to it).  Called once per handle, outside any stream
// capture.
static void allow_lds_fn(const void* fn, int bytes) {
    // a refusal only limits the tiles: a launch beyond the default LDS
    // then fails with an error status
    if (hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, bytes) != hipSuccess)
        (void)hipGetLastError();
}

template <typename R, int V, bool RV, int NT, int MODE>
static void allow_lds_mode(int bytes) {
    if constexpr (MODE >= PM_LAST)
        allow_lds_fn(reinterpret_cast<const void*>(&k_block_pass<R, V, RV, MODE, NT, 2>), bytes);
    // not instantiated (launch_f takes 512 threads there)
    constexpr bool narrow = NT == 1024 && V == 2 && sizeof(R) == 8 && MODE == PM_DOTNORM;
    if constexpr (!narrow) {
        allow_lds_fn(reinterpret_cast<const void*>(&k_block_pass<R, V, RV, MODE, NT, 3>), bytes);
        allow_lds_fn(reinterpret_cast<const void*>(&k_block_pass<R, V, RV, MODE, NT, 4>), bytes);
    }
}

template <typename R, int V, bool RV, int NT>
static void allow_lds_modes(int bytes) {
    allow_lds_mode<R, V, RV, NT, PM_NORMAL>(bytes);
    allow_lds_mode<R, V, RV, NT, PM_M1>(bytes);
    allow_lds_mode<R, V, RV, NT, PM_LAST>(bytes);
    allow_lds_mode<R, V, RV, NT, PM_NORM>(bytes);
    allow_lds_mode<R, V, RV, NT, PM_DOTNORM>(bytes);
}

template <typename R, int V, int NT>
static void allow_lds(int bytes) {
    allow_lds_modes<R, V, true, NT>(bytes);
    if constexpr (!(NT == 1024 && V == 2 && sizeof(R) == 8))   // else not instantiated
        allow_lds_modes<R, V, false, NT>(bytes);
}

void prepare_term3(int max_lds) {
    allow_lds<float, 1, 1024>(max_lds);
    allow_lds<float, 2, 1024>(max_lds);
    allow_lds<float, 2, 512>(max_lds);
    allow_lds<float, 3, 512>(max_lds);
    allow_lds<double, 1, 1024>(max_lds);
    allow_lds<double, 2, 1024>(max_lds);
    allow_lds<double, 2, 512>(max_lds);
    allow_lds<double, 3, 512>(max_lds);
}

template <typename R, int V, bool RV, int T>
static void launch_t(int mode, const PassArgs& a, const Blocking& b, hipStream_t s) {
    switch (mode) {
        case PM_LAST: launch_f<R, V, RV, PM_LAST, T>(a, b, s); return;
        case PM_NORM: launch_f<R, V, RV, PM_NORM, T>(a, b, s); return;
        case PM_DOTNORM: launch_f<R, V, RV, PM_DOTNORM, T>(a, b, s); return;
        default: break;
    }
    if constexpr (T >= 3) {
        switch (mode) {
            case PM_NORMAL: launch_f<R, V, RV, PM_NORMAL, T>(a, b, s); return;
            case PM_M1: launch_f<R, V, RV, PM_M1, T>(a, b, s); return;
            default: break;
        }
    }
    throw Error{TBG_EINVAL, "internal: unknown pass mode for this pass length"};
}

template <typename R, int V, bool RV>
static void launch_v(int mode, int T, const PassArgs& a, const Blocking& b, hipStream_t s) {
    switch (T) {
        case 2: launch_t<R, V, RV, 2>(mode, a, b, s); break;
        case 3: launch_t<R, V, RV, 3>(mode, a, b, s); break;
        case 4: launch_t<R, V, RV, 4>(mode, a, b, s); break;
        default: throw Error{TBG_EINVAL, "internal: two to four terms per pass"};
    }
}

template <typename R>
void launch_pass(int mode, int T, bool real_vals, const PassArgs& a, const Blocking& b,
                 hipStream_t s) {
    const int V = term3_vectors_per_lane(a.K);
    if (a.K > 64 || a.lgG < 0 || a.lgG > 6 || (1 << a.lgG) * V != a.K)
        throw Error{TBG_EINVAL, "internal: bad vector slot count"};
    if (b.row_bytes != a.K * (int)sizeof(typename CplxT<R>::T))
        throw Error{TBG_EINVAL, "internal: blocking built for another row size"};
    if (T < 2 || T > b.depth + 1)
        throw Error{TBG_EINVAL, "internal: more terms per pass than the tables' halo rings allow"};
    if ((mode == PM_NORM || mode == PM_DOTNORM) && (a.part == nullptr || a.nv < 3 * a.K))
        throw Error{TBG_EINVAL, "internal: epilogue without a partial-sum buffer"};
    if (mode == PM_DOTNORM && a.psi0 == nullptr)
        throw Error{TBG_EINVAL, "internal: dot epilogue without psi0"};
    if (mode < PM_LAST && (a.q1 == nullptr || a.q2 == nullptr))
        throw Error{TBG_EINVAL, "internal: a pass that continues the step needs q1 and q2"};
#define TBG_T3(VV)                                                             \
    do {                                                                       \
        if (real_vals) launch_v<R, VV, true>(mode, T, a, b, s);                \
        else launch_v<R, VV, false>(mode, T, a, b, s);                         \
    } while (0)
    switch (V) {
        case 1: TBG_T3(1); break;
        case 2: TBG_T3(2); break;
        case 3: TBG_T3(3); break;
        default: throw Error{TBG_EINVAL, "internal: unsupported vectors per lane"};
    }
#undef TBG_T3
}
template void launch_pass<float>(int, int, bool, const PassArgs&, const Blocking&, hipStream_t);
template void launch_pass<double>(int, int, bool, const PassArgs&, const Blocking&, hipStream_t);

}  // namespace tbg
