// The tile-pass kernel k_block_pass and its launch (tables in
// tbg_block.hip).  Its instances are many (real type x states per lane x
// value layout x pass mode x workgroup size x terms per pass), so each
// (real type, terms per pass) pair is instantiated in a translation unit of
// its own (and value layout, real or complex, for three to five terms),
// tbg_pass_*.hip, and the build compiles them in parallel.
#pragma once

#include <algorithm>
#include <type_traits>

#include "tbg_block.h"
#include "tbg_vec.h"

namespace tbg {

// One lane's rows l = l0, l0 + step, ... < n of a phase when every row has
// NJ <= 4 entries: the row number and the NJ table entries (index idx(j, l),
// value at vals[j * nA + l]) of the next row are loaded before the current
// row's work, so their latency hides behind it.  The last row loads its own
// entries again instead of a next row's: the loads are never conditional,
// so the compiler can count them and does not wait for them at the top of
// the next row's work.  body(l, t) does the row's work from its prefetched
// entries t (sums via tab_sum, in j order as row_sum).  With GR the row numbers come from rows[l], else they are the
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
        const RowTab<R, RV, NJ> nxt = load(more ? ln : l);
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
// says so) to keep enough gathers in flight; tiles of up to 768 rows with
// one state per lane take 4 waves, several workgroups per CU.  G = K/V
// lanes share a row, NT/G rows per pass of a phase.
//
// T terms m .. m+T-1 per pass (2 <= T <= the tables' depth + 1), phase by
// phase, n[k] = rows of R and the first k halo rings:
//   1: v_m on n[T-1] rows from the gathered v_{m-1} (HBM/L2) and v_{m-2},
//      into LDS area A; the n[T-2] rows of phase 2 also park their v_{m-1}
//      in (the still unused) area B: read next to this tile's gathers of the
//      same rows, it mostly hits L2;
//   2: v_{m+1} on n[T-2] rows from A and the parked v_{m-1}, into B;
//   3: v_{m+2} on n[T-3] rows from B and v_m.  For T >= 4 it goes over v_m
//      in A (each row reads its own v_m first, and the phase gathers only
//      from B) and the owned rows' output so far waits in a register queue;
//   4: v_{m+3} from A (v_{m+2}) and B (v_{m+1}): for T = 4 on the owned
//      rows (the last phase); for T = 5 on n[1] rows over v_{m+1} in B;
//   5: (T = 5) v_{m+4} on the owned rows from B (v_{m+3}) and A (v_{m+2}).
// The last phase works on the owned rows only: acc gets the pass's terms
// (read first, formed from v_0 for PM_M1, or the partial of phase 3), the
// last two terms go to q1, q2 unless the pass ends the step, and the
// ending modes sum the FP64 norm / dot per tile into part.
template <typename R, int V, bool RV, int MODE, int NT, int T>
__global__ void __launch_bounds__(NT) k_block_pass(PassArgs a) {
    static_assert(T >= 2 && T <= 5, "two to five terms per pass");
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

    R ar[V], ai[V], cf[T == 5 ? 7 : 6][V];   // c_{m-2} .. c_{m+3} (c_{m+4})
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
        for (int q = 0; q < (T == 5 ? 7 : 6); ++q) cf[q][v] = (R)a.c[q][cl];
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

    // phase 1: v_m.  Each row's own v_{m-1} is loaded with its gathers and
    // parked in B after them (rows beyond B's n[T-2] into the spare row at
    // nB_max): no branch separates load and store, so each row waits one
    // memory round trip, not two.
    const int spare = a.nB_max;
    phase(table_rows, m.n[T - 1], idxG, gatherG,
          [&](int, int gr) { return vld<R, V>(a.p1, (int64_t)gr * K + k0); },
          [&](int l, int gr, const VT& s, const VT& own) {
              lds_st(ldsB, l < m.n[T - 2] ? l : spare, own);
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
        } else if constexpr (T == 5) {
            // As with four terms, the owned rows' output waits in a register
            // queue (part[HELD - 1] leaves first; rows beyond HELD per lane
            // go through acc), here across phases 3 to 5: phase 4 pops each
            // queued partial, adds its term and pushes it back, so the
            // queue is in row order again for phase 5.  Every queue update
            // is a select over all HELD entries: with pops and pushes in
            // one phase, conditional moves would leave the queue in scratch.
            constexpr int HELD = V == 1 ? 1 : (V * sizeof(R) <= 8 ? 4 : 2);
            VT part[HELD];
#pragma unroll
            for (int q = 0; q < HELD; ++q) part[q] = vzero<R, V>();
            // when c: part[q] = part[q - 1] for q > 0
            auto shift_if = [&](bool c) {
#pragma unroll
                for (int q = HELD - 1; q > 0; --q) part[q] = vsel<R, V>(c, part[q - 1], part[q]);
            };
            // when c: shift, then part[0] = x
            auto shift_in = [&](bool c, VT x) {
                shift_if(c);
                part[0] = vsel<R, V>(c, x, part[0]);
            };
            // the first of held partials to part[HELD - 1]
            auto align = [&](int held) {
#pragma unroll
                for (int i = 0; i < HELD - 1; ++i) shift_if(i < HELD - held);
            };
            int held = 0;
            // phase 3: v_{m+2} on R + E1 + E2, over v_m in A
            phase(owned_rows, m.n[2], idxL, gatherB,
                  [&](int l, int gr) { return l < nR ? acc0(gr) : vzero<R, V>(); },
                  [&](int l, int gr, const VT& s, const VT& o0) {
                      const VT vm = lds(ldsA, l);
                      const VT v2 = recur<R, V, true>(ar, ai, beta, s, vm);
                      lds_st(ldsA, l, v2);
                      if (l < nR) {
                          VT o = o0;
                          axpy<R, V>(o, cf[2], vm);
                          axpy<R, V>(o, cf[3], lds(ldsB, l));
                          axpy<R, V>(o, cf[4], v2);
                          const bool push = held < HELD;
                          if (!push) vst<R, V>(a.acc, (int64_t)gr * K + k0, o);
                          shift_in(push, o);
                          held += push;
                      }
                  });
            align(held);
            __syncthreads();
            // phase 4: v_{m+3} on R + E1 over v_{m+1} in B (each row reads
            // its own v_{m+1} first; the phase gathers only from A); the
            // owned rows' v_{m+3} to q1
            int seen = 0;
            phase(owned_rows, m.n[1], idxL, gatherA,
                  [&](int l, int gr) {
                      if (l < nR && ++seen > held) return vld<R, V>(a.acc, (int64_t)gr * K + k0);
                      return vzero<R, V>();
                  },
                  [&](int l, int gr, const VT& s, const VT& o0) {
                      const VT v3 = recur<R, V, true>(ar, ai, beta, s, lds(ldsB, l));
                      lds_st(ldsB, l, v3);
                      if (l < nR) {
                          const int64_t vidx = (int64_t)gr * K + k0;
                          if constexpr (!END) vst<R, V>(a.q1, vidx, v3);
                          const bool mine = seen <= held;
                          VT o = vsel<R, V>(mine, part[HELD - 1], o0);
                          axpy<R, V>(o, cf[5], v3);
                          if (!mine) vst<R, V>(a.acc, vidx, o);
                          shift_in(mine, o);
                      }
                  });
            align(held);
            __syncthreads();
            // phase 5 (last): v_{m+4} on the owned rows from B (v_{m+3})
            // and A (v_{m+2}), the output from the queue (or acc)
            int used = 0;
            phase(owned_rows, nR, idxL, gatherB,
                  [&](int, int gr) {
                      const bool mine = used < held;
                      const VT o = part[HELD - 1];
                      shift_if(mine);
                      used += mine;
                      if (mine) return o;
                      return vld<R, V>(a.acc, (int64_t)gr * K + k0);
                  },
                  [&](int l, int gr, const VT& s, const VT& o0) {
                      const int64_t vidx = (int64_t)gr * K + k0;
                      const VT v4 = recur<R, V, true>(ar, ai, beta, s, lds(ldsA, l));
                      if constexpr (!END) vst<R, V>(a.q2, vidx, v4);
                      VT o = o0;
                      axpy<R, V>(o, cf[T + 1], v4);
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

template <typename R, int V, bool RV, int MODE, int T>
static void launch_f(const PassArgs& a, const Blocking& b, hipStream_t s) {
    // the instances term3_launch_threads can ask for; complex128 states with
    // the dot epilogue after three or more terms need more than the 128
    // registers a lane has at 16 waves, so they take 8, and so do three
    // complex128 states per lane
    constexpr bool wide =
        V == 1 || (V == 3 && sizeof(R) == 4) ||
        (V == 2 && (sizeof(R) == 4 || (RV && !(MODE == PM_DOTNORM && T >= 3))));
    int nt = term3_launch_threads((int)sizeof(R), RV, a.K, b.nA_max);
    if (!wide && V >= 2) nt = 512;
    // areas A and B and B's spare row (phase 1 parks the rows beyond B there)
    size_t lds = (size_t)(b.nA_max + b.nB_max + 1) * a.K * sizeof(typename CplxT<R>::T);
    if constexpr (MODE == PM_NORM || MODE == PM_DOTNORM)   // the waves' partial sums
        lds = std::max(lds, (size_t)(nt / 64) * 3 * a.K * sizeof(double));
    if constexpr (V == 1) {
        if (nt == 256) {
            hipLaunchKernelGGL((k_block_pass<R, V, RV, MODE, 256, T>), dim3(b.nblocks),
                               dim3(256), lds, s, a);
            TBG_HIP(hipGetLastError());
            return;
        }
    }
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
static inline void allow_lds_fn(const void* fn, int bytes) {
    // a refusal only limits the tiles: a launch beyond the default LDS
    // then fails with an error status
    if (hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, bytes) != hipSuccess)
        (void)hipGetLastError();
}

template <typename R, int V, bool RV, int NT, int T>
static void allow_lds_modes(int bytes) {
    if constexpr (T >= 3) {
        allow_lds_fn(reinterpret_cast<const void*>(&k_block_pass<R, V, RV, PM_NORMAL, NT, T>),
                     bytes);
        allow_lds_fn(reinterpret_cast<const void*>(&k_block_pass<R, V, RV, PM_M1, NT, T>), bytes);
    }
    allow_lds_fn(reinterpret_cast<const void*>(&k_block_pass<R, V, RV, PM_LAST, NT, T>), bytes);
    allow_lds_fn(reinterpret_cast<const void*>(&k_block_pass<R, V, RV, PM_NORM, NT, T>), bytes);
    // not instantiated (launch_f takes 512 threads there)
    constexpr bool narrow = NT == 1024 && V == 2 && sizeof(R) == 8 && T >= 3;
    if constexpr (!narrow)
        allow_lds_fn(reinterpret_cast<const void*>(&k_block_pass<R, V, RV, PM_DOTNORM, NT, T>),
                     bytes);
}

template <typename R, int V, int NT, int T, bool RV>
static void allow_lds_v(int bytes) {
    if constexpr (RV || !(NT == 1024 && V == 2 && sizeof(R) == 8))   // else not instantiated
        allow_lds_modes<R, V, RV, NT, T>(bytes);
}

template <typename R, int T, bool RV>
void allow_lds_rt(int bytes) {
    allow_lds_v<R, 1, 256, T, RV>(bytes);
    allow_lds_v<R, 1, 1024, T, RV>(bytes);
    allow_lds_v<R, 2, 1024, T, RV>(bytes);
    allow_lds_v<R, 2, 512, T, RV>(bytes);
    if constexpr (sizeof(R) == 4) allow_lds_v<R, 3, 1024, T, RV>(bytes);
    allow_lds_v<R, 3, 512, T, RV>(bytes);
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

template <typename R, int T, bool RV>
void launch_pass_rt(int mode, const PassArgs& a, const Blocking& b, hipStream_t s) {
    switch (term3_vectors_per_lane(a.K)) {
        case 1: launch_t<R, 1, RV, T>(mode, a, b, s); break;
        case 2: launch_t<R, 2, RV, T>(mode, a, b, s); break;
        case 3: launch_t<R, 3, RV, T>(mode, a, b, s); break;
        default: throw Error{TBG_EINVAL, "internal: unsupported vectors per lane"};
    }
}

}  // namespace tbg
