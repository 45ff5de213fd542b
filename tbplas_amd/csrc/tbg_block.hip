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
    TBG_REQUIRE(depth >= 2 && depth <= 4, "internal: halo depth must be 2 to 4");
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
    int64_t target = std::max<int64_t>(
        64, (int64_t)(lds / ((depth == 2 ? 2.6 : depth == 3 ? 3.0 : 3.4) * row_bytes)));
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
    int64_t sched[4] = {0, 0, 0, 0};
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
        for (int q = 0; q < 4; ++q) sched[q] = 0;
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
                size_t ends[5];
                ends[0] = rows.size();
                for (int k = 1; k <= depth; ++k) {   // E1 .. E_depth
                    ring(k == 1 ? base : ends[k - 2], ends[k - 1]);
                    ends[k] = rows.size();
                }
                const int nA = (int)(ends[depth] - base), nB = (int)(ends[depth - 1] - base);
                // (LDS positions are 16 bits: nA < 65536 whenever A fits the LDS)
                // (A, B and B's spare row)
                if ((int64_t)(nA + nB + 1) * row_bytes > budget || nA > 65535) {
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
                for (int k = 0; k < 5; ++k) m.n[k] = (int)(ends[std::min(k, depth)] - base);
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
                // distinct rows of v_{m-1} the first phase gathers: A and
                // the ring beyond its last one
                int64_t outer = 0;
                for (int l = nB; l < nA; ++l) {
                    const int r = rows[base + l];
                    for (int jj = 0, w = width(r); jj < w; ++jj) {
                        const int c = col(r, jj);
                        if (mark[c] != b) {
                            mark[c] = b;
                            ++outer;
                        }
                    }
                }
                sched[0] += (int64_t)wA * nA * (int64_t)(sizeof(int) + vsz) +
                            (int64_t)wA * nB * (int64_t)sizeof(uint16_t) +
                            (int64_t)nA * (int64_t)sizeof(int) + (int64_t)sizeof(BlockMeta);
                sched[1] += nA + outer;
                sched[2] += nA;
                sched[3] += r1 - r0;
                nA_max = std::max(nA_max, nA);
                nB_max = std::max(nB_max, nB);
            }
        }
        seg_block.push_back((int)meta.size());
        if ((int64_t)(nA_max + nB_max + 1) * row_bytes <= h->max_lds) break;
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
    for (int q = 0; q < 4; ++q) B->sched[q] = sched[q];
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
    // one state per lane on small tiles: 4-wave workgroups, several per CU,
    // so that a launch over a small sample runs its tiles at once
    if (V == 1) return nA <= 768 ? 256 : 1024;
    if (V == 3 && elem_bytes == 4) return (int64_t)nA * (K / V) >= 2048 ? 1024 : 512;
    if (V != 2 || (elem_bytes == 8 && !real_vals)) return 512;
    return (int64_t)nA * (K / V) >= 2048 ? 1024 : 512;
}

template <typename R, int T>
static void allow_lds_t(int max_lds) {
    allow_lds_rt<R, T, true>(max_lds);
    allow_lds_rt<R, T, false>(max_lds);
}

void prepare_term3(int max_lds) {
    allow_lds_t<float, 2>(max_lds);
    allow_lds_t<float, 3>(max_lds);
    allow_lds_t<float, 4>(max_lds);
    allow_lds_t<double, 2>(max_lds);
    allow_lds_t<double, 3>(max_lds);
    allow_lds_t<double, 4>(max_lds);
    allow_lds_t<float, 5>(max_lds);
    allow_lds_t<double, 5>(max_lds);
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
#define TBG_PASS_T(TT)                                                         \
    do {                                                                       \
        if (real_vals) launch_pass_rt<R, TT, true>(mode, a, b, s);             \
        else launch_pass_rt<R, TT, false>(mode, a, b, s);                      \
    } while (0)
    switch (T) {
        case 2: TBG_PASS_T(2); break;
        case 3: TBG_PASS_T(3); break;
        case 4: TBG_PASS_T(4); break;
        case 5: TBG_PASS_T(5); break;
        default: throw Error{TBG_EINVAL, "internal: two to five terms per pass"};
    }
#undef TBG_PASS_T
}
template void launch_pass<float>(int, int, bool, const PassArgs&, const Blocking&, hipStream_t);
template void launch_pass<double>(int, int, bool, const PassArgs&, const Blocking&, hipStream_t);

}  // namespace tbg
