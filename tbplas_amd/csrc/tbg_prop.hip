// Propagation kernels for the TBPM hot path on CDNA4 (gfx950).
//
// Layout (HBM):
//   * Hamiltonian in sliced ELL: 64-row slices, each stored column-major
//     (element j of local row r at soff[s] + j*64 + r), width = longest row
//     of the slice, padding entries (col = own row, value 0).  One wavefront
//     owns one slice, so width and offset are wave-uniform scalar loads and
//     column/value loads are fully coalesced.
//   * K state vectors interleaved per row: element (row, k) at row*K + k,
//     complex<R> with R = float or double.  A lane owns V consecutive vectors
//     of one row (vectors_per_lane: 1 to 4), and G = K/V lanes share a row,
//     so each gather x[col] is one contiguous K*sizeof(complex) segment read
//     by G lanes; a wave covers its slice in G passes.
//
// The fused term kernel computes v_m = alpha H v_{m-1} + beta v_{m-2} for all
// K vectors of a block of slices, writes v_m over v_{m-2} in place, and every
// third term folds v_{m-2}, v_{m-1}, v_m into the step output, so the output
// vector is read and written once per three terms instead of every term
// (the reference makes one amxpy and one axpy pass per term,
// tbplas/fortran/propagation.f90:42-48).  The last term of a step also forms
// the FP64 correlation <psi0|psi(t)> and |psi(t)|^2 per vector
// (tbplas/fortran/tbpm.f90:81-86, funcs.f90:175-192) as per-block partial
// sums, reduced in a fixed order by k_reduce (deterministic, no atomics).
#include <type_traits>

#include "tbg_prop.h"
#include "tbg_vec.h"

namespace tbg {


// STAGE: the slice's own 64 rows of v_{m-1} are first read once, coalesced,
// into LDS; gathers whose column falls inside the slice (for lattice models
// most of them: the +-1 neighbours) and the v_{m-1} term of the
// accumulation are then served from LDS, and only the remaining gathers go
// through L2.  Each wave owns its own 64*K-element LDS area.
//
// RV: the values are stored as R (every imaginary part of H is zero), which
// halves the value bytes of the sweep; the products then skip the terms
// that would multiply an exact zero, so the sums are the same as with the
// complex values.
template <typename R, int V, int FL, bool STAGE, bool RV>
__global__ void __launch_bounds__(256) k_term(TermArgs a) {
    using C = typename CplxT<R>::T;
    extern __shared__ __align__(16) unsigned char tbg_smem[];
    const int lane = threadIdx.x & 63;
    const int wv = threadIdx.x >> 6;
    const int c = remap_chunk(blockIdx.x, gridDim.x);
    if (a.counter != nullptr && blockIdx.x == 0 && threadIdx.x == 0) *a.counter += 1;

    const int lgG = a.lgG;
    const int G = 1 << lgG;
    const int g = lane & (G - 1);
    const int rsub = lane >> lgG;
    const int rpp = 64 >> lgG;
    const int K = a.K;
    const int k0 = g * V;

    R ar[V], ai[V], cp[V], cm[V], cc[V];
#pragma unroll
    for (int v = 0; v < V; ++v) {
        const int k = k0 + v;
        const int cl = (int)((a.cls_mask >> k) & 1ull);
        ar[v] = (R)a.ar[cl];
        ai[v] = ((a.bwd_mask >> k) & 1ull) ? (R)(-a.ai[cl]) : (R)a.ai[cl];
        cp[v] = (R)a.cp[cl];
        cm[v] = (R)a.cm[cl];
        cc[v] = (R)a.cc[cl];
    }
    const R beta = (R)a.beta;

    double dre[V], dim_[V], nrm[V];
#pragma unroll
    for (int v = 0; v < V; ++v) { dre[v] = 0.0; dim_[v] = 0.0; nrm[v] = 0.0; }

    unsigned char* stage = tbg_smem + (size_t)wv * 64 * K * sizeof(C);
    const int64_t s0 = a.chunk[c], s1 = a.chunk[c + 1];
    for (int64_t sl = s0 + wv; sl < s1; sl += 4) {
        const int64_t off = a.soff[sl];
        const int w = a.sw[sl];
        const int* colp = a.cols + off;
        const void* valp = reinterpret_cast<const unsigned char*>(a.vals) +
                           off * (RV ? sizeof(R) : sizeof(C));
        const int64_t row0 = sl << 6;
        if constexpr (STAGE) {
            for (int p = 0; p < G; ++p) {
                const int lr = p * rpp + rsub;
                vst<R, V>(stage, (int64_t)lr * K + k0,
                          vld<R, V>(a.pin, (row0 + lr) * (int64_t)K + k0));
            }
            __builtin_amdgcn_wave_barrier();
        }
        for (int p = 0; p < G; ++p) {
            const int lr = p * rpp + rsub;
            const int64_t vidx = (row0 + lr) * (int64_t)K + k0;
            Vec<R, V> prev, mid, accv;
            if constexpr ((FL & F_PREV) != 0) prev = stream_ld<R, V>(a.pprev, vidx);
            if constexpr ((FL & F_MID) != 0) {
                if constexpr (STAGE) mid = vld<R, V>(stage, (int64_t)lr * K + k0);
                else mid = vld<R, V>(a.pin, vidx);
            }
            if constexpr ((FL & F_ACCRMW) != 0) accv = stream_ld<R, V>(a.acc, vidx);
            const auto idx = [&](int j) { return colp[j * 64 + lr]; };
            const auto ent = [&](int j) { return j * 64 + lr; };
            const auto gather = [&](int col) {
                if constexpr (STAGE) {
                    const uint32_t d = (uint32_t)((int64_t)col - row0);
                    if (d < 64u) return vld<R, V>(stage, (int64_t)d * K + k0);
                }
                return vld<R, V>(a.pin, (int64_t)col * K + k0);
            };
            const Vec<R, V> s = row_sum<R, V, RV>(w, valp, idx, ent, gather);
            const Vec<R, V> cur = recur<R, V, (FL & F_PREV) != 0>(ar, ai, beta, s, prev);
            if constexpr ((FL & F_STORE) != 0) stream_st<R, V>(a.pout, vidx, cur);
            if constexpr ((FL & F_ACC) != 0) {
                Vec<R, V> o;
                if constexpr ((FL & F_ACCRMW) != 0) o = accv; else o = vzero<R, V>();
                if constexpr ((FL & F_ACCPREV) != 0) axpy<R, V>(o, cp, prev);
                if constexpr ((FL & F_MID) != 0) axpy<R, V>(o, cm, mid);
                axpy<R, V>(o, cc, cur);
                stream_st<R, V>(a.acc, vidx, o);
                if constexpr ((FL & F_DOT) != 0) {
                    const Vec<R, V> p0 = stream_ld<R, V>(a.psi0, vidx);
#pragma unroll
                    for (int v = 0; v < V; ++v) {
                        const double pr = p0.re[v], pi = p0.im[v];
                        const double orr = o.re[v], oi = o.im[v];
                        dre[v] = fma(pr, orr, fma(pi, oi, dre[v]));
                        dim_[v] = fma(pr, oi, fma(-pi, orr, dim_[v]));
                    }
                }
                if constexpr ((FL & F_NORM) != 0) {
#pragma unroll
                    for (int v = 0; v < V; ++v) {
                        const double orr = o.re[v], oi = o.im[v];
                        nrm[v] = fma(orr, orr, fma(oi, oi, nrm[v]));
                    }
                }
            }
        }
        if constexpr (STAGE) __builtin_amdgcn_wave_barrier();
    }

    if constexpr ((FL & (F_DOT | F_NORM)) != 0) {
        // Lanes with equal g hold the same vectors: butterfly over them.
        for (int o = G; o < 64; o <<= 1) {
#pragma unroll
            for (int v = 0; v < V; ++v) {
                dre[v] += __shfl_xor(dre[v], o);
                dim_[v] += __shfl_xor(dim_[v], o);
                nrm[v] += __shfl_xor(nrm[v], o);
            }
        }
        __shared__ double red[4][192];
        if (rsub == 0) {
#pragma unroll
            for (int v = 0; v < V; ++v) {
                const int k = k0 + v;
                red[wv][k * 3 + 0] = dre[v];
                red[wv][k * 3 + 1] = dim_[v];
                red[wv][k * 3 + 2] = nrm[v];
            }
        }
        __syncthreads();
        for (int q = threadIdx.x; q < 3 * K; q += 256)
            a.part[(int64_t)c * a.nv + q] = ((red[0][q] + red[1][q]) + red[2][q]) + red[3][q];
    }
}

// Fixed-order reduction of partial rows: block (q, inst) sums value q over
// the instance's chunks (strided per thread, then an LDS tree).
__global__ void __launch_bounds__(256)
k_reduce(const double* part, int nv, const int* inst_chunk, double* out,
         int64_t inst_stride, const int* counter, int toff) {
    __shared__ double sh[256];
    const int q = blockIdx.x, inst = blockIdx.y;
    const int c0 = inst_chunk[inst], c1 = inst_chunk[inst + 1];
    double acc = 0.0;
    for (int c = c0 + threadIdx.x; c < c1; c += 256) acc += part[(int64_t)c * nv + q];
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const int t = (counter ? *counter : 0) + toff;
        out[inst * inst_stride + (int64_t)t * nv + q] = sh[0];
    }
}

// Thread i handles layout row r = i % n of vector k = i / n; it reads or
// writes the caller's row perm[r] (r without a permutation).
template <typename R>
__global__ void k_pack(const double2* __restrict__ src, int K, int64_t n,
                       typename CplxT<R>::T* __restrict__ dst, int Kp, int slot0,
                       const int* __restrict__ perm) {
    // consecutive threads take consecutive states of one row: the stores of
    // a row's slots are one contiguous run
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)K * n) return;
    const int64_t r = i / K, k = i - r * K;
    const int64_t c = perm ? (int64_t)perm[r] : r;
    const double2 v = src[k * n + c];
    typename CplxT<R>::T o;
    o.x = (R)v.x;
    o.y = (R)v.y;
    dst[r * Kp + slot0 + k] = o;
}

template <typename R>
__global__ void k_unpack(const typename CplxT<R>::T* __restrict__ src, int Kp, int slot0,
                         int64_t n, double2* __restrict__ dst, int K, const int* __restrict__ perm) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)K * n) return;
    const int64_t r = i / K, k = i - r * K;
    const int64_t c = perm ? (int64_t)perm[r] : r;
    const typename CplxT<R>::T v = src[r * Kp + slot0 + k];
    dst[k * n + c] = make_double2((double)v.x, (double)v.y);
}

// conj(a) . b and |b|^2 per vector over the first K slots, per-chunk
// partials.  Thread t takes slot t % Kp of every (256 / Kp)-th row from row
// t / Kp, so a wave reads whole rows (coalesced); the row groups' sums are
// then added in group order.  Kp <= 256.
template <typename R>
__global__ void __launch_bounds__(256)
k_dotnorm(const int64_t* chunk, const typename CplxT<R>::T* a, const typename CplxT<R>::T* b,
          int Kp, int K, double* part) {
    const int c = remap_chunk(blockIdx.x, gridDim.x);
    const int64_t r0 = chunk[c] * 64, r1 = chunk[c + 1] * 64;
    __shared__ double sh[256 * 3];
    const int RG = 256 / Kp;
    const int k = threadIdx.x % Kp, rg = threadIdx.x / Kp;
    double dr = 0, di = 0, nn = 0;
    if (rg < RG && k < K) {
        for (int64_t r = r0 + rg; r < r1; r += RG) {
            const auto x = a[r * Kp + k];
            const auto y = b[r * Kp + k];
            const double xr = x.x, xi = x.y, yr = y.x, yi = y.y;
            dr = fma(xr, yr, fma(xi, yi, dr));
            di = fma(xr, yi, fma(-xi, yr, di));
            nn = fma(yr, yr, fma(yi, yi, nn));
        }
    }
    sh[threadIdx.x * 3 + 0] = dr;
    sh[threadIdx.x * 3 + 1] = di;
    sh[threadIdx.x * 3 + 2] = nn;
    __syncthreads();
    for (int q = threadIdx.x; q < 3 * K; q += 256) {
        const int kk = q / 3, e = q - kk * 3;
        double t = 0.0;
        for (int g = 0; g < RG; ++g) t += sh[(g * Kp + kk) * 3 + e];
        part[(int64_t)c * 3 * K + q] = t;
    }
}

// y[r][sy+k] = sum_j M[r,j] x[col][sx+k], k < nk, one thread per (row, k).
template <typename R>
__global__ void k_spmv(const int* __restrict__ cols, const typename CplxT<R>::T* __restrict__ mv,
                       const int64_t* __restrict__ soff, const int* __restrict__ sw,
                       int64_t n_pad, const typename CplxT<R>::T* __restrict__ x, int Kx, int sx,
                       typename CplxT<R>::T* __restrict__ y, int Ky, int sy, int nk) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_pad * nk) return;
    const int64_t r = i / nk;
    const int k = (int)(i - r * nk);
    const int64_t sl = r >> 6;
    const int lr = (int)(r & 63);
    const int64_t off = soff[sl];
    const int w = sw[sl];
    R sr = 0, si = 0;
    for (int j = 0; j < w; ++j) {
        const int col = cols[off + j * 64 + lr];
        const auto h = mv[off + j * 64 + lr];
        const auto v = x[(int64_t)col * Kx + sx + k];
        sr = fma(h.x, v.x, sr);
        sr = fma(-h.y, v.y, sr);
        si = fma(h.x, v.y, si);
        si = fma(h.y, v.x, si);
    }
    typename CplxT<R>::T o;
    o.x = sr;
    o.y = si;
    y[r * Ky + sy + k] = o;
}

__global__ void k_fill_zero(uint4* p, int64_t n16) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = make_uint4(0, 0, 0, 0);
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------
#define TBG_TERM_FLAG_LIST(X)                                                  \
    X(F_STORE)                                                                 \
    X(F_PREV | F_STORE)                                                        \
    X(F_PREV | F_STORE | F_ACC | F_ACCPREV | F_MID)                            \
    X(F_PREV | F_STORE | F_ACC | F_ACCRMW | F_ACCPREV | F_MID)                 \
    X(F_PREV | F_ACC | F_ACCPREV | F_MID)                                      \
    X(F_PREV | F_ACC | F_ACCPREV | F_MID | F_NORM)                             \
    X(F_PREV | F_ACC | F_ACCPREV | F_MID | F_NORM | F_DOT)                     \
    X(F_PREV | F_ACC | F_ACCRMW | F_ACCPREV | F_MID)                           \
    X(F_PREV | F_ACC | F_ACCRMW | F_ACCPREV | F_MID | F_NORM)                  \
    X(F_PREV | F_ACC | F_ACCRMW | F_ACCPREV | F_MID | F_NORM | F_DOT)          \
    X(F_PREV | F_ACC | F_ACCRMW)                                               \
    X(F_PREV | F_STORE | F_ACC | F_ACCRMW)                                     \
    X(F_PREV | F_ACC | F_ACCRMW | F_NORM)                                      \
    X(F_PREV | F_ACC | F_ACCRMW | F_NORM | F_DOT)                              \
    X(F_PREV | F_ACC | F_ACCRMW | F_MID)                                       \
    X(F_PREV | F_ACC | F_ACCRMW | F_MID | F_NORM)                              \
    X(F_PREV | F_ACC | F_ACCRMW | F_MID | F_NORM | F_DOT)

template <typename R, int V, bool STAGE, bool RV>
static void launch_term_v(int flags, const TermArgs& a, int nchunks, hipStream_t s) {
    const size_t lds = STAGE ? (size_t)4 * 64 * a.K * sizeof(typename CplxT<R>::T) : 0;
    switch (flags) {
#define TBG_CASE(F)                                                            \
    case (F):                                                                  \
        hipLaunchKernelGGL((k_term<R, V, (F), STAGE, RV>), dim3(nchunks), dim3(256), lds, s, a); \
        break;
        TBG_TERM_FLAG_LIST(TBG_CASE)
#undef TBG_CASE
        default:
            throw Error{TBG_EINVAL, "internal: unsupported term flag set " + std::to_string(flags)};
    }
    TBG_HIP(hipGetLastError());
}

template <typename R>
void launch_term(int flags, const TermArgs& a, int nchunks, hipStream_t s) {
    const int V = vectors_per_lane<R>(a.K);
    if (a.K > 64 || a.lgG < 0 || a.lgG > 6 || (1 << a.lgG) * V != a.K)
        throw Error{TBG_EINVAL, "internal: bad vector slot count"};
    // Stage the slice in LDS while one wave's area stays within 4 KB
    // (complex64 K <= 8, complex128 K <= 4): 16 KB per block.
    const bool stage = a.stage && 64 * a.K * sizeof(typename CplxT<R>::T) <= 4096;
    const bool rv = a.real_vals != 0;
#define TBG_TERM_V(VV)                                                         \
    do {                                                                       \
        if (stage) {                                                           \
            if (rv) launch_term_v<R, VV, true, true>(flags, a, nchunks, s);    \
            else launch_term_v<R, VV, true, false>(flags, a, nchunks, s);      \
        } else {                                                               \
            if (rv) launch_term_v<R, VV, false, true>(flags, a, nchunks, s);   \
            else launch_term_v<R, VV, false, false>(flags, a, nchunks, s);     \
        }                                                                      \
    } while (0)
    switch (V) {
        case 1: TBG_TERM_V(1); break;
        case 2: TBG_TERM_V(2); break;
        case 3: TBG_TERM_V(3); break;
        case 4: TBG_TERM_V(4); break;
        default:
            throw Error{TBG_EINVAL, "internal: unsupported vectors per lane"};
    }
#undef TBG_TERM_V
}
template void launch_term<float>(int, const TermArgs&, int, hipStream_t);
template void launch_term<double>(int, const TermArgs&, int, hipStream_t);

void launch_reduce(const double* part, int nv, const int* inst_chunk, int ninst, double* out,
                   int64_t inst_stride, const int* counter, int toff, hipStream_t s) {
    hipLaunchKernelGGL(k_reduce, dim3(nv, ninst), dim3(256), 0, s, part, nv, inst_chunk, out,
                       inst_stride, counter, toff);
    TBG_HIP(hipGetLastError());
}

template <typename R>
void launch_pack(const double* src, int K, int64_t n, void* dst, int Kp, int slot0,
                 const int* perm, hipStream_t s) {
    const int64_t tot = (int64_t)K * n;
    if (tot == 0) return;
    hipLaunchKernelGGL(k_pack<R>, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s,
                       reinterpret_cast<const double2*>(src), K, n,
                       reinterpret_cast<typename CplxT<R>::T*>(dst), Kp, slot0, perm);
    TBG_HIP(hipGetLastError());
}
template void launch_pack<float>(const double*, int, int64_t, void*, int, int, const int*, hipStream_t);
template void launch_pack<double>(const double*, int, int64_t, void*, int, int, const int*, hipStream_t);

template <typename R>
void launch_unpack(const void* src, int Kp, int slot0, int64_t n, double* dst, int K,
                   const int* perm, hipStream_t s) {
    const int64_t tot = (int64_t)K * n;
    if (tot == 0) return;
    hipLaunchKernelGGL(k_unpack<R>, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s,
                       reinterpret_cast<const typename CplxT<R>::T*>(src), Kp, slot0, n,
                       reinterpret_cast<double2*>(dst), K, perm);
    TBG_HIP(hipGetLastError());
}
template void launch_unpack<float>(const void*, int, int, int64_t, double*, int, const int*, hipStream_t);
template void launch_unpack<double>(const void*, int, int, int64_t, double*, int, const int*, hipStream_t);

template <typename R>
void launch_dotnorm(const Plan& p, const void* a, const void* b, int Kp, int K, double* part,
                    hipStream_t s) {
    if (Kp < 1 || Kp > 256 || K > Kp) throw Error{TBG_EINVAL, "internal: bad slot count"};
    hipLaunchKernelGGL(k_dotnorm<R>, dim3(p.nchunks), dim3(256), 0, s, p.chunk.as<int64_t>(),
                       reinterpret_cast<const typename CplxT<R>::T*>(a),
                       reinterpret_cast<const typename CplxT<R>::T*>(b), Kp, K, part);
    TBG_HIP(hipGetLastError());
}
template void launch_dotnorm<float>(const Plan&, const void*, const void*, int, int, double*, hipStream_t);
template void launch_dotnorm<double>(const Plan&, const void*, const void*, int, int, double*, hipStream_t);

template <typename R>
void launch_spmv(const tbg_ham* H, const void* mvals, const void* x, int Kx, int sx, void* y,
                 int Ky, int sy, int nk, hipStream_t s) {
    const int64_t n_pad = H->nslices * 64;
    const int64_t tot = n_pad * nk;
    if (tot == 0) return;
    hipLaunchKernelGGL(k_spmv<R>, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s,
                       H->cols.as<int>(), reinterpret_cast<const typename CplxT<R>::T*>(mvals),
                       H->slice_off.as<int64_t>(), H->slice_w.as<int>(), n_pad,
                       reinterpret_cast<const typename CplxT<R>::T*>(x), Kx, sx,
                       reinterpret_cast<typename CplxT<R>::T*>(y), Ky, sy, nk);
    TBG_HIP(hipGetLastError());
}
template void launch_spmv<float>(const tbg_ham*, const void*, const void*, int, int, void*, int, int, int, hipStream_t);
template void launch_spmv<double>(const tbg_ham*, const void*, const void*, int, int, void*, int, int, int, hipStream_t);

void launch_fill_zero(void* p, size_t bytes, hipStream_t s) {
    if (bytes == 0) return;
    if (bytes % 16) throw Error{TBG_EINVAL, "internal: fill size not a multiple of 16"};
    const int64_t n16 = (int64_t)(bytes / 16);
    const int64_t blocks = std::min<int64_t>((n16 + 255) / 256, 8192);
    hipLaunchKernelGGL(k_fill_zero, dim3((unsigned)blocks), dim3(256), 0, s,
                       reinterpret_cast<uint4*>(p), n16);
    TBG_HIP(hipGetLastError());
}

void make_plan(tbg_handle* h, Plan& plan, const std::vector<int64_t>& inst_slices,
               int target_chunks) {
    const int ninst = (int)inst_slices.size() - 1;
    plan.nslices = inst_slices.back();
    plan.chunk_host.clear();
    plan.inst_chunk.assign(1, 0);
    const int64_t total = std::max<int64_t>(plan.nslices, 1);
    // Chunks hold a multiple of 4 slices (one per wave) where possible.
    for (int i = 0; i < ninst; ++i) {
        const int64_t a = inst_slices[i], b = inst_slices[i + 1];
        const int64_t ns = b - a;
        int64_t want = std::max<int64_t>(1, (int64_t)((double)target_chunks * ns / total + 0.5));
        int64_t per = (ns + want - 1) / want;
        per = std::max<int64_t>(4, (per + 3) / 4 * 4);
        for (int64_t s = a; s < b; s += per) plan.chunk_host.push_back(s);
        if (ns == 0) plan.chunk_host.push_back(a);
        plan.inst_chunk.push_back((int)plan.chunk_host.size());
    }
    plan.chunk_host.push_back(plan.nslices);
    plan.nchunks = (int)plan.chunk_host.size() - 1;
    plan.chunk.alloc(plan.chunk_host.size() * sizeof(int64_t));
    plan.inst_chunk_dev.alloc(plan.inst_chunk.size() * sizeof(int));
    TBG_HIP(hipMemcpyAsync(plan.chunk.p, plan.chunk_host.data(),
                           plan.chunk_host.size() * sizeof(int64_t), hipMemcpyHostToDevice,
                           h->stream));
    TBG_HIP(hipMemcpyAsync(plan.inst_chunk_dev.p, plan.inst_chunk.data(),
                           plan.inst_chunk.size() * sizeof(int), hipMemcpyHostToDevice,
                           h->stream));
    TBG_HIP(hipStreamSynchronize(h->stream));
}

}  // namespace tbg
