// CSR assembly, rescale estimate and the sliced-ELL device layout.
//
// Assembly places every entry of the emission stream of the reference
// builders by row with integer atomics (counts are exact), stores its
// emission key, and then restores the exact insertion order by sorting each
// row's keys: the result is deterministic and bit-identical to the serial
// builders whatever order the atomics resolved in.
//   key -1        : the diagonal entry of the row
//   key 2*ih      : hop ih as (hop_i, hop_j, hop_v, +dr)
//   key 2*ih + 1  : hop ih as (hop_j, hop_i, conj(hop_v), -dr)
// (tbplas/cython/sample.pyx:148-257 fills each row in exactly this order.)
#include <algorithm>

#include "tbg_csr.h"

namespace tbg {

// ---------------------------------------------------------------------------
// Prefix sums (int64), three phases: block sums, scan of sums, apply.
// ---------------------------------------------------------------------------
constexpr int kScanTile = 1024;  // 256 threads x 4 elements

__device__ __forceinline__ int64_t block_scan_incl(int64_t v, int64_t* sh_wave) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    for (int o = 1; o < 64; o <<= 1) {
        const int64_t t = __shfl_up(v, o);
        if (lane >= o) v += t;
    }
    if (lane == 63) sh_wave[wv] = v;
    __syncthreads();
    int64_t add = 0;
    for (int w = 0; w < wv; ++w) add += sh_wave[w];
    __syncthreads();
    return v + add;
}

__global__ void __launch_bounds__(256) k_scan_sums(const int64_t* in, int64_t n, int64_t* bsum) {
    __shared__ int64_t shw[4];
    const int64_t base = (int64_t)blockIdx.x * kScanTile + threadIdx.x * 4;
    int64_t s = 0;
    for (int i = 0; i < 4; ++i)
        if (base + i < n) s += in[base + i];
    const int64_t incl = block_scan_incl(s, shw);
    if (threadIdx.x == 255) bsum[blockIdx.x] = incl;
}

// Exclusive scan of a (in place) in one block, tile by tile with a carry.
__global__ void __launch_bounds__(256) k_scan_single(int64_t* a, int64_t n) {
    __shared__ int64_t shw[4];
    __shared__ int64_t carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (int64_t t0 = 0; t0 < n; t0 += kScanTile) {
        const int64_t base = t0 + threadIdx.x * 4;
        int64_t v[4], s = 0;
        for (int i = 0; i < 4; ++i) {
            v[i] = (base + i < n) ? a[base + i] : 0;
            s += v[i];
        }
        const int64_t incl = block_scan_incl(s, shw);
        int64_t run = carry + incl - s;
        for (int i = 0; i < 4; ++i) {
            if (base + i < n) a[base + i] = run;
            run += v[i];
        }
        __syncthreads();
        if (threadIdx.x == 255) carry += incl;
        __syncthreads();
    }
}

__global__ void __launch_bounds__(256)
k_scan_apply(const int64_t* in, int64_t n, const int64_t* boff, int64_t* out, int exclusive) {
    __shared__ int64_t shw[4];
    const int64_t base = (int64_t)blockIdx.x * kScanTile + threadIdx.x * 4;
    int64_t v[4], s = 0;
    for (int i = 0; i < 4; ++i) {
        v[i] = (base + i < n) ? in[base + i] : 0;
        s += v[i];
    }
    const int64_t incl = block_scan_incl(s, shw);
    int64_t run = boff[blockIdx.x] + incl - s;
    for (int i = 0; i < 4; ++i) {
        if (exclusive) {
            if (base + i < n) out[base + i] = run;
            run += v[i];
        } else {
            run += v[i];
            if (base + i < n) out[base + i] = run;
        }
    }
}

void scan_i64(tbg_handle* h, const int64_t* in, int64_t n, int64_t* out, bool exclusive,
              DevBuf& tmp) {
    if (n == 0) return;
    const int64_t nb = (n + kScanTile - 1) / kScanTile;
    tmp.ensure((size_t)nb * sizeof(int64_t));
    hipLaunchKernelGGL(k_scan_sums, dim3((unsigned)nb), dim3(256), 0, h->stream, in, n,
                       tmp.as<int64_t>());
    TBG_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_scan_single, dim3(1), dim3(256), 0, h->stream, tmp.as<int64_t>(), nb);
    TBG_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_scan_apply, dim3((unsigned)nb), dim3(256), 0, h->stream, in, n,
                       tmp.as<int64_t>(), out, exclusive ? 1 : 0);
    TBG_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// Placement of the emission stream
// ---------------------------------------------------------------------------
// diag_mode: 0 none, 1 |orb_eng| >= cutoff (build_ham_dxy), 2 orb_eng != 0
__device__ __forceinline__ bool diag_kept(int mode, double e, double cutoff) {
    return mode == 1 ? (fabs(e) >= cutoff) : (mode == 2 ? (e != 0.0) : false);
}

__global__ void k_count(int64_t n, int64_t m, const double* orb_eng, int diag_mode, double cutoff,
                        const int64_t* hop_i, const int64_t* hop_j, unsigned long long* cnt1) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && diag_kept(diag_mode, orb_eng[i], cutoff)) atomicAdd(&cnt1[i + 1], 1ull);
    if (i < m) {
        atomicAdd(&cnt1[hop_i[i] + 1], 1ull);
        atomicAdd(&cnt1[hop_j[i] + 1], 1ull);
    }
}

__global__ void k_place(int64_t n, int64_t m, const double* orb_eng, int diag_mode, double cutoff,
                        const int64_t* hop_i, const int64_t* hop_j, const int64_t* indptr,
                        unsigned int* cursor, int64_t* keys) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && diag_kept(diag_mode, orb_eng[i], cutoff)) {
        const unsigned int c = atomicAdd(&cursor[i], 1u);
        keys[indptr[i] + c] = -1;
    }
    if (i < m) {
        const int64_t a = hop_i[i], b = hop_j[i];
        const unsigned int ca = atomicAdd(&cursor[a], 1u);
        keys[indptr[a] + ca] = 2 * i;
        const unsigned int cb = atomicAdd(&cursor[b], 1u);
        keys[indptr[b] + cb] = 2 * i + 1;
    }
}

__device__ __forceinline__ void sort_keys(int64_t* k, int64_t len) {
    for (int64_t a = 1; a < len; ++a) {
        const int64_t v = k[a];
        int64_t b = a - 1;
        while (b >= 0 && k[b] > v) { k[b + 1] = k[b]; --b; }
        k[b + 1] = v;
    }
}

// Insertion order (build_ham_dxy_fast): sort keys, then materialise.
__global__ void k_dxy_fill(int64_t n, const double* orb_eng, const int64_t* hop_i,
                           const int64_t* hop_j, const double2* hop_v, const double* dr,
                           const int64_t* indptr, int64_t* indices, double2* hop, double* dx,
                           double* dy) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const int64_t p0 = indptr[r], p1 = indptr[r + 1];
    sort_keys(indices + p0, p1 - p0);
    for (int64_t p = p0; p < p1; ++p) {
        const int64_t key = indices[p];
        if (key < 0) {
            indices[p] = r;
            hop[p] = make_double2(orb_eng[r], 0.0);
            dx[p] = 0.0;
            dy[p] = 0.0;
        } else {
            const int64_t ih = key >> 1;
            const double2 v = hop_v[ih];
            if ((key & 1) == 0) {
                indices[p] = hop_j[ih];
                hop[p] = v;
                dx[p] = dr[ih * 3 + 0];
                dy[p] = dr[ih * 3 + 1];
            } else {
                indices[p] = hop_i[ih];
                hop[p] = make_double2(v.x, -v.y);
                dx[p] = -dr[ih * 3 + 0];
                dy[p] = -dr[ih * 3 + 1];
            }
        }
    }
}

// Canonical order (scipy dia + csr + csr^H): per row, sort the entries by
// (column, group, hop index) with group 0 = diagonal, 1 = hop_i row,
// 2 = hop_j row; sum each column as (g0 + g1) + g2 with every group summed
// in hop order; drop exact zeros.  Merged entries are written back in place
// (cols in keys, values in vals) and the merged count to cnt[r].
__device__ __forceinline__ int64_t canon_col(int64_t key, int64_t r, const int64_t* hop_i,
                                             const int64_t* hop_j) {
    if (key < 0) return r;
    return (key & 1) ? hop_i[key >> 1] : hop_j[key >> 1];
}

__device__ __forceinline__ bool canon_less(int64_t ka, int64_t kb, int64_t r, const int64_t* hi,
                                           const int64_t* hj) {
    const int64_t ca = canon_col(ka, r, hi, hj), cb = canon_col(kb, r, hi, hj);
    if (ca != cb) return ca < cb;
    const int ga = ka < 0 ? 0 : 1 + (int)(ka & 1), gb = kb < 0 ? 0 : 1 + (int)(kb & 1);
    if (ga != gb) return ga < gb;
    return ka < kb;
}

__global__ void k_canon_merge(int64_t n, const double* orb_eng, const int64_t* hop_i,
                              const int64_t* hop_j, const double2* hop_v, const int64_t* tptr,
                              int64_t* keys, double2* vals, int64_t* cnt) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const int64_t p0 = tptr[r], len = tptr[r + 1] - p0;
    int64_t* k = keys + p0;
    for (int64_t a = 1; a < len; ++a) {
        const int64_t v = k[a];
        int64_t b = a - 1;
        while (b >= 0 && canon_less(v, k[b], r, hop_i, hop_j)) { k[b + 1] = k[b]; --b; }
        k[b + 1] = v;
    }
    int64_t out = 0;
    int64_t p = 0;
    while (p < len) {
        const int64_t col = canon_col(k[p], r, hop_i, hop_j);
        double2 gs[3];
        bool has[3] = {false, false, false};
        for (int q = 0; q < 3; ++q) gs[q] = make_double2(0.0, 0.0);
        while (p < len && canon_col(k[p], r, hop_i, hop_j) == col) {
            const int64_t key = k[p];
            int g;
            double2 v;
            if (key < 0) { g = 0; v = make_double2(orb_eng[r], 0.0); }
            else if ((key & 1) == 0) { g = 1; v = hop_v[key >> 1]; }
            else { g = 2; v = hop_v[key >> 1]; }
            if (!has[g]) { gs[g] = v; has[g] = true; }
            else { gs[g].x += v.x; gs[g].y += v.y; }
            ++p;
        }
        // half^H: conjugate of the summed half entry
        if (has[2]) gs[2].y = -gs[2].y;
        double2 tot;
        // scipy adds a missing operand as 0 (op(x, 0) = x + 0)
        tot.x = (gs[0].x + gs[1].x) + gs[2].x;
        tot.y = (gs[0].y + gs[1].y) + gs[2].y;
        if (tot.x != 0.0 || tot.y != 0.0) {
            k[out] = col;
            vals[p0 + out] = tot;
            ++out;
        }
    }
    cnt[r + 1] = out;
}

__global__ void k_canon_compact(int64_t n, const int64_t* tptr, const int64_t* keys,
                                const double2* vals, const int64_t* indptr, int64_t* indices,
                                double2* data) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const int64_t s = tptr[r], d = indptr[r], len = indptr[r + 1] - d;
    for (int64_t q = 0; q < len; ++q) {
        indices[d + q] = keys[s + q];
        data[d + q] = vals[s + q];
    }
}

// get_rescale (sample.pyx:75-114): orb_eng[r] + |hop| in emission order.
__global__ void k_rowabs(int64_t n, const double* orb_eng, const double2* hop_v,
                         const int64_t* tptr, int64_t* keys, double* rowsum) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const int64_t p0 = tptr[r], p1 = tptr[r + 1];
    sort_keys(keys + p0, p1 - p0);
    double s = orb_eng[r];
    for (int64_t p = p0; p < p1; ++p) {
        const double2 v = hop_v[keys[p] >> 1];
        s = s + sqrt(v.x * v.x + v.y * v.y);
    }
    rowsum[r] = s;
}

__global__ void k_max_partial(const double* x, int64_t n, double* part) {
    __shared__ double sh[256];
    double m = -INFINITY;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
        m = fmax(m, x[i]);
    sh[threadIdx.x] = m;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

// ---------------------------------------------------------------------------
// Host helpers
// ---------------------------------------------------------------------------
static unsigned blocks_for(int64_t n, int bs = 256) {
    return (unsigned)std::max<int64_t>(1, (n + bs - 1) / bs);
}

static void check_hops(int64_t n, int64_t m) {
    TBG_REQUIRE(n > 0, "number of orbitals must be positive");
    TBG_REQUIRE(m >= 0, "number of hops must be non-negative");
    TBG_REQUIRE(m < (int64_t(1) << 61), "too many hops");
}

__global__ void k_check_hops(int64_t m, const int64_t* hop_i, const int64_t* hop_j, int64_t n,
                             int* bad) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < m && (hop_i[i] < 0 || hop_i[i] >= n || hop_j[i] < 0 || hop_j[i] >= n))
        atomicOr(bad, 1);
}

// Every hop index must lie in [0, n) before any kernel scatters by it.
static void check_hop_indices(tbg_handle* h, int64_t n, int64_t m, const int64_t* hop_i,
                              const int64_t* hop_j) {
    if (m == 0) return;
    DevBuf bad(sizeof(int));
    TBG_HIP(hipMemsetAsync(bad.p, 0, sizeof(int), h->stream));
    hipLaunchKernelGGL(k_check_hops, dim3(blocks_for(m)), dim3(256), 0, h->stream, m, hop_i,
                       hop_j, n, bad.as<int>());
    TBG_HIP(hipGetLastError());
    int hb = 0;
    TBG_HIP(hipMemcpyAsync(&hb, bad.p, sizeof(int), hipMemcpyDeviceToHost, h->stream));
    TBG_HIP(hipStreamSynchronize(h->stream));
    TBG_REQUIRE(hb == 0, "hop_i/hop_j index out of range [0, n)");
}

// indptr (n+1) from the emission stream; returns nnz.
int64_t place_count(tbg_handle* h, int64_t n, int64_t m, const double* orb_eng, int diag_mode,
                    double cutoff, const int64_t* hop_i, const int64_t* hop_j, int64_t* indptr,
                    DevBuf& tmp) {
    check_hops(n, m);
    check_hop_indices(h, n, m, hop_i, hop_j);
    DevBuf cnt((size_t)(n + 1) * sizeof(int64_t));
    TBG_HIP(hipMemsetAsync(cnt.p, 0, (size_t)(n + 1) * sizeof(int64_t), h->stream));
    const int64_t mx = std::max(n, m);
    hipLaunchKernelGGL(k_count, dim3(blocks_for(mx)), dim3(256), 0, h->stream, n, m, orb_eng,
                       diag_mode, cutoff, hop_i, hop_j, cnt.as<unsigned long long>());
    TBG_HIP(hipGetLastError());
    scan_i64(h, cnt.as<int64_t>(), n + 1, indptr, false, tmp);
    int64_t nnz = 0;
    TBG_HIP(hipMemcpyAsync(&nnz, indptr + n, sizeof(int64_t), hipMemcpyDeviceToHost, h->stream));
    TBG_HIP(hipStreamSynchronize(h->stream));
    return nnz;
}

void place_keys(tbg_handle* h, int64_t n, int64_t m, const double* orb_eng, int diag_mode,
                double cutoff, const int64_t* hop_i, const int64_t* hop_j, const int64_t* indptr,
                int64_t* keys) {
    DevBuf cursor((size_t)n * sizeof(unsigned int));
    TBG_HIP(hipMemsetAsync(cursor.p, 0, (size_t)n * sizeof(unsigned int), h->stream));
    const int64_t mx = std::max(n, m);
    hipLaunchKernelGGL(k_place, dim3(blocks_for(mx)), dim3(256), 0, h->stream, n, m, orb_eng,
                       diag_mode, cutoff, hop_i, hop_j, indptr, cursor.as<unsigned int>(), keys);
    TBG_HIP(hipGetLastError());
}

void dxy_fill(tbg_handle* h, int64_t n, int64_t m, const double* orb_eng, const int64_t* hop_i,
              const int64_t* hop_j, const double* hop_v, const double* dr, double cutoff,
              const int64_t* indptr, int64_t* indices, double* hop, double* dx, double* dy) {
    check_hops(n, m);
    check_hop_indices(h, n, m, hop_i, hop_j);
    // The caller's indptr must match the stream (it comes from the count
    // call); verify before scattering into the caller's arrays.
    {
        DevBuf tmp;
        DevBuf ref((size_t)(n + 1) * sizeof(int64_t));
        place_count(h, n, m, orb_eng, 1, cutoff, hop_i, hop_j, ref.as<int64_t>(), tmp);
        std::vector<int64_t> a(n + 1), b(n + 1);
        TBG_HIP(hipMemcpyAsync(a.data(), ref.p, (n + 1) * sizeof(int64_t), hipMemcpyDeviceToHost,
                               h->stream));
        TBG_HIP(hipMemcpyAsync(b.data(), indptr, (n + 1) * sizeof(int64_t),
                               hipMemcpyDeviceToHost, h->stream));
        TBG_HIP(hipStreamSynchronize(h->stream));
        TBG_REQUIRE(a == b, "indptr does not match the hop list (call tbg_ham_dxy_count first)");
    }
    place_keys(h, n, m, orb_eng, 1, cutoff, hop_i, hop_j, indptr, indices);
    hipLaunchKernelGGL(k_dxy_fill, dim3(blocks_for(n)), dim3(256), 0, h->stream, n, orb_eng,
                       hop_i, hop_j, reinterpret_cast<const double2*>(hop_v), dr, indptr, indices,
                       reinterpret_cast<double2*>(hop), dx, dy);
    TBG_HIP(hipGetLastError());
    TBG_HIP(hipStreamSynchronize(h->stream));
}

// Canonical CSR: merged per-row entries in temporaries; returns nnz and the
// merged-count prefix in indptr.  When indices/data are non-null, compacts.
int64_t canon_build(tbg_handle* h, int64_t n, int64_t m, const double* orb_eng,
                    const int64_t* hop_i, const int64_t* hop_j, const double* hop_v,
                    int64_t* indptr_out, const int64_t* indptr_in, int64_t* indices,
                    double* data) {
    DevBuf tmp;
    DevBuf tptr((size_t)(n + 1) * sizeof(int64_t));
    const int64_t ntmp = place_count(h, n, m, orb_eng, 2, 0.0, hop_i, hop_j, tptr.as<int64_t>(), tmp);
    DevBuf keys((size_t)std::max<int64_t>(ntmp, 1) * sizeof(int64_t));
    DevBuf vals((size_t)std::max<int64_t>(ntmp, 1) * sizeof(double2));
    place_keys(h, n, m, orb_eng, 2, 0.0, hop_i, hop_j, tptr.as<int64_t>(), keys.as<int64_t>());
    DevBuf cnt((size_t)(n + 1) * sizeof(int64_t));
    TBG_HIP(hipMemsetAsync(cnt.p, 0, sizeof(int64_t), h->stream));
    hipLaunchKernelGGL(k_canon_merge, dim3(blocks_for(n)), dim3(256), 0, h->stream, n, orb_eng,
                       hop_i, hop_j, reinterpret_cast<const double2*>(hop_v), tptr.as<int64_t>(),
                       keys.as<int64_t>(), vals.as<double2>(), cnt.as<int64_t>());
    TBG_HIP(hipGetLastError());
    int64_t nnz = 0;
    DevBuf own;
    int64_t* ip = indptr_out;
    if (!ip) {
        own.alloc((size_t)(n + 1) * sizeof(int64_t));
        ip = own.as<int64_t>();
    }
    scan_i64(h, cnt.as<int64_t>(), n + 1, ip, false, tmp);
    TBG_HIP(hipMemcpyAsync(&nnz, ip + n, sizeof(int64_t), hipMemcpyDeviceToHost, h->stream));
    if (indices) {
        std::vector<int64_t> a(n + 1), b(n + 1);
        TBG_HIP(hipMemcpyAsync(a.data(), ip, (n + 1) * sizeof(int64_t), hipMemcpyDeviceToHost,
                               h->stream));
        TBG_HIP(hipMemcpyAsync(b.data(), indptr_in, (n + 1) * sizeof(int64_t),
                               hipMemcpyDeviceToHost, h->stream));
        TBG_HIP(hipStreamSynchronize(h->stream));
        TBG_REQUIRE(a == b, "indptr does not match the hop list (call tbg_ham_csr_count first)");
        hipLaunchKernelGGL(k_canon_compact, dim3(blocks_for(n)), dim3(256), 0, h->stream, n,
                           tptr.as<int64_t>(), keys.as<int64_t>(), vals.as<double2>(), indptr_in,
                           indices, reinterpret_cast<double2*>(data));
        TBG_HIP(hipGetLastError());
    }
    TBG_HIP(hipStreamSynchronize(h->stream));
    return nnz;
}

double get_rescale(tbg_handle* h, int64_t n, int64_t m, const double* orb_eng,
                   const int64_t* hop_i, const int64_t* hop_j, const double* hop_v) {
    DevBuf tmp;
    DevBuf tptr((size_t)(n + 1) * sizeof(int64_t));
    const int64_t ntmp = place_count(h, n, m, orb_eng, 0, 0.0, hop_i, hop_j, tptr.as<int64_t>(), tmp);
    DevBuf keys((size_t)std::max<int64_t>(ntmp, 1) * sizeof(int64_t));
    place_keys(h, n, m, orb_eng, 0, 0.0, hop_i, hop_j, tptr.as<int64_t>(), keys.as<int64_t>());
    DevBuf rs((size_t)n * sizeof(double));
    hipLaunchKernelGGL(k_rowabs, dim3(blocks_for(n)), dim3(256), 0, h->stream, n, orb_eng,
                       reinterpret_cast<const double2*>(hop_v), tptr.as<int64_t>(),
                       keys.as<int64_t>(), rs.as<double>());
    TBG_HIP(hipGetLastError());
    const int nb = (int)std::min<int64_t>(1024, (n + 255) / 256);
    DevBuf part((size_t)nb * sizeof(double));
    hipLaunchKernelGGL(k_max_partial, dim3(nb), dim3(256), 0, h->stream, rs.as<double>(), n,
                       part.as<double>());
    TBG_HIP(hipGetLastError());
    std::vector<double> hp(nb);
    TBG_HIP(hipMemcpyAsync(hp.data(), part.p, nb * sizeof(double), hipMemcpyDeviceToHost,
                           h->stream));
    TBG_HIP(hipStreamSynchronize(h->stream));
    double mx = -INFINITY;
    for (double v : hp) mx = std::max(mx, v);
    return mx;
}

// ---------------------------------------------------------------------------
// Sliced-ELL layout for propagation
// ---------------------------------------------------------------------------
// perm (optional, new row -> caller row) orders the rows of the layout;
// inv is its inverse.  Without perm both are the identity.
__global__ void k_invert_perm(int64_t n, const int64_t* perm, int* inv, int* bad) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int64_t p = perm[i];
    if (p < 0 || p >= n) {
        atomicOr(bad, 1);
        return;
    }
    if (atomicExch(&inv[p], (int)i) != -1) atomicOr(bad, 2);
}

__global__ void k_perm_i32(int64_t n, const int64_t* perm, int* out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = (int)perm[i];
}

__global__ void k_slice_width(int64_t n, int64_t nslices, const int64_t* indptr, const int* perm,
                              int* sw, int64_t* selems) {
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nslices) return;
    int64_t w = 0;
    for (int r = 0; r < 64; ++r) {
        const int64_t row = s * 64 + r;
        if (row < n) {
            const int64_t src = perm ? perm[row] : row;
            w = max(w, indptr[src + 1] - indptr[src]);
        }
    }
    sw[s] = (int)w;
    selems[s] = w * 64;
}

// REAL: every imaginary part of H is zero; values are stored as R.
template <typename R, bool REAL>
__global__ void k_sell_fill(int64_t n, int64_t nslices, const int64_t* indptr,
                            const int64_t* indices, const double2* hop, const double* dx,
                            const double* dy, double cur_scale, const int* perm, const int* inv,
                            const int64_t* soff, const int* sw, int* cols, void* vals_,
                            typename CplxT<R>::T* jx, typename CplxT<R>::T* jy) {
    const int64_t row = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= nslices * 64) return;
    const int64_t s = row >> 6;
    const int lr = (int)(row & 63);
    const int64_t off = soff[s];
    const int w = sw[s];
    // entries keep the caller's order within the row (same sums); only the
    // row position and the column numbers follow the permutation
    const int64_t src = row < n ? (perm ? (int64_t)perm[row] : row) : 0;
    const int64_t p0 = row < n ? indptr[src] : 0;
    const int64_t len = row < n ? indptr[src + 1] - p0 : 0;
    for (int j = 0; j < w; ++j) {
        const int64_t e = off + (int64_t)j * 64 + lr;
        typename CplxT<R>::T v, x, y;
        v.x = v.y = x.x = x.y = y.x = y.y = R(0);
        int col = (int)(row < n ? row : 0);
        if (j < len) {
            const int64_t p = p0 + j;
            col = inv ? inv[indices[p]] : (int)indices[p];
            const double2 hv = hop[p];
            v.x = (R)hv.x;
            v.y = (R)hv.y;
            if (jx) {
                // current_coefficient (funcs.f90:12-33): (0, s) * hop * d
                const double ar = -cur_scale * hv.y, ai = cur_scale * hv.x;
                x.x = (R)(ar * dx[p]);
                x.y = (R)(ai * dx[p]);
                y.x = (R)(ar * dy[p]);
                y.y = (R)(ai * dy[p]);
            }
        }
        cols[e] = col;
        if constexpr (REAL) reinterpret_cast<R*>(vals_)[e] = v.x;
        else reinterpret_cast<typename CplxT<R>::T*>(vals_)[e] = v;
        if (jx) {
            jx[e] = x;
            jy[e] = y;
        }
    }
}

__global__ void k_any_imag(int64_t nnz, const double2* hop, int* flag) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nnz && hop[i].y != 0.0) atomicOr(flag, 1);
}

__global__ void k_check_cols(int64_t nnz, const int64_t* indices, int64_t n, int* bad) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nnz && (indices[i] < 0 || indices[i] >= n)) atomicOr(bad, 1);
}

__global__ void k_check_indptr(int64_t n, const int64_t* indptr, int* bad) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r < n && indptr[r + 1] < indptr[r]) atomicOr(bad, 2);
}

void build_sell(tbg_handle* h, tbg_ham* H, const int64_t* indptr, const int64_t* indices,
                const double* hop, const double* dx, const double* dy, double cur_scale,
                const int64_t* perm) {
    const int64_t n = H->n;
    const int64_t ns = (n + 63) / 64;
    H->nslices = ns;
    int64_t nnz = 0, first = 0;
    TBG_HIP(hipMemcpyAsync(&nnz, indptr + n, sizeof(int64_t), hipMemcpyDeviceToHost, h->stream));
    TBG_HIP(hipMemcpyAsync(&first, indptr, sizeof(int64_t), hipMemcpyDeviceToHost, h->stream));
    TBG_HIP(hipStreamSynchronize(h->stream));
    TBG_REQUIRE(first == 0, "indptr[0] must be 0");
    TBG_REQUIRE(nnz >= 0, "indptr[n] must be non-negative");
    H->nnz = nnz;
    {
        DevBuf bad(sizeof(int));
        TBG_HIP(hipMemsetAsync(bad.p, 0, sizeof(int), h->stream));
        hipLaunchKernelGGL(k_check_indptr, dim3(blocks_for(n)), dim3(256), 0, h->stream, n,
                           indptr, bad.as<int>());
        TBG_HIP(hipGetLastError());
        int hb0 = 0;
        TBG_HIP(hipMemcpyAsync(&hb0, bad.p, sizeof(int), hipMemcpyDeviceToHost, h->stream));
        TBG_HIP(hipStreamSynchronize(h->stream));
        TBG_REQUIRE(hb0 == 0, "indptr must be non-decreasing");
        if (nnz > 0) {
            hipLaunchKernelGGL(k_check_cols, dim3(blocks_for(nnz)), dim3(256), 0, h->stream, nnz,
                               indices, n, bad.as<int>());
            TBG_HIP(hipGetLastError());
        }
        int hb = 0;
        TBG_HIP(hipMemcpyAsync(&hb, bad.p, sizeof(int), hipMemcpyDeviceToHost, h->stream));
        TBG_HIP(hipStreamSynchronize(h->stream));
        TBG_REQUIRE(hb == 0, "column index out of range [0, n)");
    }
    DevBuf inv;
    H->ordered = perm != nullptr;
    if (H->ordered) {
        TBG_REQUIRE(n > 0, "a row order needs at least one row");
        H->perm.alloc((size_t)n * sizeof(int));
        inv.alloc((size_t)n * sizeof(int));
        DevBuf bad(sizeof(int));
        TBG_HIP(hipMemsetAsync(inv.p, 0xff, (size_t)n * sizeof(int), h->stream));
        TBG_HIP(hipMemsetAsync(bad.p, 0, sizeof(int), h->stream));
        hipLaunchKernelGGL(k_invert_perm, dim3(blocks_for(n)), dim3(256), 0, h->stream, n, perm,
                           inv.as<int>(), bad.as<int>());
        TBG_HIP(hipGetLastError());
        int hb = 0;
        TBG_HIP(hipMemcpyAsync(&hb, bad.p, sizeof(int), hipMemcpyDeviceToHost, h->stream));
        TBG_HIP(hipStreamSynchronize(h->stream));
        TBG_REQUIRE(hb == 0, "row order is not a permutation of [0, n)");
        hipLaunchKernelGGL(k_perm_i32, dim3(blocks_for(n)), dim3(256), 0, h->stream, n, perm,
                           H->perm.as<int>());
        TBG_HIP(hipGetLastError());
    }
    const int* pm = H->ordered ? H->perm.as<int>() : nullptr;
    const int* iv = H->ordered ? inv.as<int>() : nullptr;
    H->slice_w.alloc((size_t)ns * sizeof(int));
    DevBuf selems((size_t)ns * sizeof(int64_t));
    hipLaunchKernelGGL(k_slice_width, dim3(blocks_for(ns)), dim3(256), 0, h->stream, n, ns, indptr,
                       pm, H->slice_w.as<int>(), selems.as<int64_t>());
    TBG_HIP(hipGetLastError());
    H->slice_off.alloc((size_t)(ns + 1) * sizeof(int64_t));
    DevBuf tmp;
    TBG_HIP(hipMemsetAsync(H->slice_off.p, 0, sizeof(int64_t), h->stream));
    scan_i64(h, selems.as<int64_t>(), ns, H->slice_off.as<int64_t>() + 1, false, tmp);
    int64_t padded = 0;
    TBG_HIP(hipMemcpyAsync(&padded, H->slice_off.as<int64_t>() + ns, sizeof(int64_t),
                           hipMemcpyDeviceToHost, h->stream));
    std::vector<int> hw(ns);
    TBG_HIP(hipMemcpyAsync(hw.data(), H->slice_w.p, ns * sizeof(int), hipMemcpyDeviceToHost,
                           h->stream));
    TBG_HIP(hipStreamSynchronize(h->stream));
    H->padded = padded;
    H->max_width = 0;
    for (int w : hw) H->max_width = std::max(H->max_width, w);
    {
        DevBuf flag(sizeof(int));
        TBG_HIP(hipMemsetAsync(flag.p, 0, sizeof(int), h->stream));
        if (nnz > 0) {
            hipLaunchKernelGGL(k_any_imag, dim3(blocks_for(nnz)), dim3(256), 0, h->stream, nnz,
                               reinterpret_cast<const double2*>(hop), flag.as<int>());
            TBG_HIP(hipGetLastError());
        }
        int hf = 0;
        TBG_HIP(hipMemcpyAsync(&hf, flag.p, sizeof(int), hipMemcpyDeviceToHost, h->stream));
        TBG_HIP(hipStreamSynchronize(h->stream));
        H->real_vals = h->real_layout && hf == 0;
    }
    const bool f32 = H->precision == TBG_FP32;
    const size_t cb = f32 ? sizeof(float2) : sizeof(double2);
    const size_t vb = H->real_vals ? cb / 2 : cb;
    H->cols.alloc((size_t)std::max<int64_t>(padded, 1) * sizeof(int));
    H->vals.alloc((size_t)std::max<int64_t>(padded, 1) * vb);
    H->has_current = dx != nullptr && dy != nullptr;
    if (H->has_current) {
        H->jx.alloc((size_t)std::max<int64_t>(padded, 1) * cb);
        H->jy.alloc((size_t)std::max<int64_t>(padded, 1) * cb);
    }
    const unsigned nb = blocks_for(ns * 64);
#define TBG_FILL(R, REAL)                                                                  \
    hipLaunchKernelGGL((k_sell_fill<R, REAL>), dim3(nb), dim3(256), 0, h->stream, n, ns,     \
                       indptr, indices, reinterpret_cast<const double2*>(hop), dx, dy,       \
                       cur_scale, pm, iv, H->slice_off.as<int64_t>(), H->slice_w.as<int>(),  \
                       H->cols.as<int>(), H->vals.p,                                         \
                       H->has_current ? H->jx.as<typename CplxT<R>::T>() : nullptr,          \
                       H->has_current ? H->jy.as<typename CplxT<R>::T>() : nullptr)
    if (f32) {
        if (H->real_vals) TBG_FILL(float, true);
        else TBG_FILL(float, false);
    } else {
        if (H->real_vals) TBG_FILL(double, true);
        else TBG_FILL(double, false);
    }
#undef TBG_FILL
    TBG_HIP(hipGetLastError());
    TBG_HIP(hipStreamSynchronize(h->stream));
}

}  // namespace tbg
