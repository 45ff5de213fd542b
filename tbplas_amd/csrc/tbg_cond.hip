// AC and DC conductivity correlation drivers and their kernels, plus the
// block-diagonal concatenation used by batched runs.
//
// AC (tbplas/fortran/tbpm.f90:193-351): the 3K states J_x psi0, J_y psi0 and
// psi0 of K samples are propagated together as one SpMM; the Fermi and
// one-minus-Fermi polynomials run in the same launches through per-vector
// coefficient classes.  After each step one kernel forms the four current
// correlations <psi2| J_a psi1_b> with the current operators stored on the
// Hamiltonian's structure.
//
// DC (tbpm.f90:497-718): phase 1 propagates psi(+t) and psi(-t) together
// (per-vector sign of the time step) and accumulates the quasi-eigenstates in
// FP64 in the same pass as the DOS correlation; phase 2 propagates J_a psi0
// and correlates with J_a |QE_i>/|<psi0|QE_i>|, which the reference
// recomputes every step (tbpm.f90:677-697) but is time independent and is
// formed once here.
#include <cmath>

#include "tbg_cond.h"

namespace tbg {

static unsigned nblk(int64_t n, int bs = 256) {
    return (unsigned)std::max<int64_t>(1, (n + bs - 1) / bs);
}

// ---------------------------------------------------------------------------
// Batch concatenation
// ---------------------------------------------------------------------------
__global__ void k_add_i64(const int64_t* src, int64_t n, int64_t add, int64_t* dst) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i] + add;
}
__global__ void k_add_i32(const int* src, int64_t n, int add, int* dst) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i] + add;
}

// Real values widened to complex (imaginary part 0) for a batch that mixes
// real and complex Hamiltonians.
template <typename R>
__global__ void k_real_to_cplx(const R* src, int64_t n, typename CplxT<R>::T* dst) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        typename CplxT<R>::T o;
        o.x = src[i];
        o.y = R(0);
        dst[i] = o;
    }
}

void concat_hams(tbg_handle* h, int n, const tbg_ham* const* hams, Concat& out) {
    int64_t ns = 0, ne = 0;
    bool all_real = true;
    for (int i = 0; i < n; ++i) {
        ns += hams[i]->nslices;
        ne += hams[i]->padded;
        all_real = all_real && hams[i]->real_vals;
    }
    TBG_REQUIRE(ns * 64 < (int64_t(1) << 31), "batch too large for 32-bit columns");
    const bool f32 = hams[0]->precision == TBG_FP32;
    const size_t cb = f32 ? sizeof(float2) : sizeof(double2);
    const size_t vb = all_real ? cb / 2 : cb;
    out.real_vals = all_real ? 1 : 0;
    out.nslices = ns;
    out.cols.alloc((size_t)std::max<int64_t>(ne, 1) * sizeof(int));
    out.vals.alloc((size_t)std::max<int64_t>(ne, 1) * vb);
    out.soff.alloc((size_t)(ns + 1) * sizeof(int64_t));
    out.sw.alloc((size_t)std::max<int64_t>(ns, 1) * sizeof(int));
    int64_t s0 = 0, e0 = 0;
    for (int i = 0; i < n; ++i) {
        const tbg_ham* H = hams[i];
        if (H->padded > 0) {
            hipLaunchKernelGGL(k_add_i32, dim3(nblk(H->padded)), dim3(256), 0, h->stream,
                               H->cols.as<int>(), H->padded, (int)(s0 * 64),
                               out.cols.as<int>() + e0);
            TBG_HIP(hipGetLastError());
            char* dst = reinterpret_cast<char*>(out.vals.p) + e0 * vb;
            if (H->real_vals && !all_real) {
                if (f32)
                    hipLaunchKernelGGL(k_real_to_cplx<float>, dim3(nblk(H->padded)), dim3(256), 0,
                                       h->stream, H->vals.as<float>(), H->padded,
                                       reinterpret_cast<float2*>(dst));
                else
                    hipLaunchKernelGGL(k_real_to_cplx<double>, dim3(nblk(H->padded)), dim3(256), 0,
                                       h->stream, H->vals.as<double>(), H->padded,
                                       reinterpret_cast<double2*>(dst));
                TBG_HIP(hipGetLastError());
            } else {
                TBG_HIP(hipMemcpyAsync(dst, H->vals.p, (size_t)H->padded * vb,
                                       hipMemcpyDeviceToDevice, h->stream));
            }
        }
        hipLaunchKernelGGL(k_add_i64, dim3(nblk(H->nslices + 1)), dim3(256), 0, h->stream,
                           H->slice_off.as<int64_t>(), H->nslices + 1, e0,
                           out.soff.as<int64_t>() + s0);
        TBG_HIP(hipGetLastError());
        TBG_HIP(hipMemcpyAsync(out.sw.as<int>() + s0, H->slice_w.p, (size_t)H->nslices * sizeof(int),
                               hipMemcpyDeviceToDevice, h->stream));
        s0 += H->nslices;
        e0 += H->padded;
    }
    TBG_HIP(hipStreamSynchronize(h->stream));
}

// ---------------------------------------------------------------------------
// Kernels
// ---------------------------------------------------------------------------
// Four current correlations per sample s: slots s (psi1_x), K+s (psi1_y),
// 2K+s (psi2).  part[c][s][q][re, im], q = xx, xy, yx, yy (tbpm.f90:284-291).
template <typename R>
__global__ void __launch_bounds__(256)
k_accorr(const int* __restrict__ cols, const typename CplxT<R>::T* __restrict__ jx,
         const typename CplxT<R>::T* __restrict__ jy, const int64_t* __restrict__ soff,
         const int* __restrict__ sw, const int64_t* __restrict__ chunk,
         const typename CplxT<R>::T* __restrict__ X, int Kp, int K, double* __restrict__ part) {
    __shared__ double sh[256 * 8];
    const int c = remap_chunk(blockIdx.x, gridDim.x);
    const int64_t r0 = chunk[c] * 64, r1 = chunk[c + 1] * 64;
    const int ng = 256 / K;
    const int s = threadIdx.x % K, grp = threadIdx.x / K;
    double d[8];
    for (int q = 0; q < 8; ++q) d[q] = 0.0;
    if (grp < ng) {
        for (int64_t r = r0 + grp; r < r1; r += ng) {
            const int64_t sl = r >> 6;
            const int lr = (int)(r & 63);
            const int64_t off = soff[sl];
            const int w = sw[sl];
            R w0r = 0, w0i = 0, w1r = 0, w1i = 0, w2r = 0, w2i = 0, w3r = 0, w3i = 0;
            for (int j = 0; j < w; ++j) {
                const int64_t e = off + (int64_t)j * 64 + lr;
                const int col = cols[e];
                const auto a = jx[e];
                const auto b = jy[e];
                const auto u = X[(int64_t)col * Kp + s];
                const auto v = X[(int64_t)col * Kp + K + s];
                w0r = fma(a.x, u.x, fma(-a.y, u.y, w0r)); w0i = fma(a.x, u.y, fma(a.y, u.x, w0i));
                w1r = fma(b.x, u.x, fma(-b.y, u.y, w1r)); w1i = fma(b.x, u.y, fma(b.y, u.x, w1i));
                w2r = fma(a.x, v.x, fma(-a.y, v.y, w2r)); w2i = fma(a.x, v.y, fma(a.y, v.x, w2i));
                w3r = fma(b.x, v.x, fma(-b.y, v.y, w3r)); w3i = fma(b.x, v.y, fma(b.y, v.x, w3i));
            }
            const auto p = X[r * Kp + 2 * K + s];
            const double pr = p.x, pi = p.y;
            const double wr[4] = {w0r, w1r, w2r, w3r}, wi[4] = {w0i, w1i, w2i, w3i};
            for (int q = 0; q < 4; ++q) {
                d[2 * q + 0] = fma(pr, wr[q], fma(pi, wi[q], d[2 * q + 0]));
                d[2 * q + 1] = fma(pr, wi[q], fma(-pi, wr[q], d[2 * q + 1]));
            }
        }
    }
    for (int q = 0; q < 8; ++q) sh[threadIdx.x * 8 + q] = d[q];
    __syncthreads();
    if (threadIdx.x < K * 8) {
        const int s2 = threadIdx.x / 8, q = threadIdx.x % 8;
        double acc = 0.0;
        for (int g = 0; g < ng; ++g) acc += sh[(g * K + s2) * 8 + q];
        part[(int64_t)c * 8 * K + s2 * 8 + q] = acc;
    }
}

// QE(r, e, s) initialised to psi0 (FP64, tbpm.f90:600-603).
// psi0 is in the caller's row order; layout row r holds caller row perm[r].
__global__ void k_qe_init(const double2* __restrict__ psi0, int64_t n, int K, int n_en,
                          int64_t n_pad, const int* __restrict__ perm, double2* __restrict__ qe) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t nq = (int64_t)n_en * K;
    if (i >= n_pad * nq) return;
    const int64_t r = i / nq;
    const int q = (int)(i - r * nq);
    const int s = q % K;
    const int64_t c = r < n ? (perm ? (int64_t)perm[r] : r) : 0;
    qe[i] = r < n ? psi0[(int64_t)s * n + c] : make_double2(0.0, 0.0);
}

// After step t (counter = t): DOS correlation <psi0|psi(+t)> and the
// quasi-eigenstate update (tbpm.f90:616-638).
template <typename R>
__global__ void __launch_bounds__(256)
k_qe(const int64_t* __restrict__ chunk, const typename CplxT<R>::T* __restrict__ X, int Kp,
     int K, const typename CplxT<R>::T* __restrict__ psi, double2* __restrict__ qe, int n_en,
     const double2* __restrict__ ptab, const double* __restrict__ wtab,
     const int* __restrict__ counter, double* __restrict__ part) {
    __shared__ double sh[256 * 2];
    const int c = remap_chunk(blockIdx.x, gridDim.x);
    const int64_t r0 = chunk[c] * 64, r1 = chunk[c + 1] * 64;
    const int t = *counter;
    const double W = wtab[t - 1];
    const double2* P = ptab + (int64_t)(t - 1) * n_en;
    const int ng = 256 / K;
    const int s = threadIdx.x % K, grp = threadIdx.x / K;
    const int64_t nq = (int64_t)n_en * K;
    double dr = 0.0, di = 0.0;
    if (grp < ng) {
        for (int64_t r = r0 + grp; r < r1; r += ng) {
            const auto p = X[r * Kp + s];
            const auto q = X[r * Kp + K + s];
            const auto z = psi[r * Kp + s];
            const double pr = p.x, pi = p.y, qr = q.x, qi = q.y;
            dr = fma((double)z.x, pr, fma((double)z.y, pi, dr));
            di = fma((double)z.x, pi, fma(-(double)z.y, pr, di));
            double2* row = qe + r * nq + s;
            for (int e = 0; e < n_en; ++e) {
                const double2 Pe = P[e];
                double2 v = row[(int64_t)e * K];
                // QE = QE + P*pos*W; QE = QE + conj(P)*neg*W
                const double ar = Pe.x * pr - Pe.y * pi, ai = Pe.x * pi + Pe.y * pr;
                v.x = v.x + ar * W;
                v.y = v.y + ai * W;
                const double br = Pe.x * qr + Pe.y * qi, bi = Pe.x * qi - Pe.y * qr;
                v.x = v.x + br * W;
                v.y = v.y + bi * W;
                row[(int64_t)e * K] = v;
            }
        }
    }
    sh[threadIdx.x * 2 + 0] = dr;
    sh[threadIdx.x * 2 + 1] = di;
    __syncthreads();
    if (threadIdx.x < 2 * K) {
        const int s2 = threadIdx.x / 2, q = threadIdx.x % 2;
        double acc = 0.0;
        for (int g = 0; g < ng; ++g) acc += sh[(g * K + s2) * 2 + q];
        part[(int64_t)c * 2 * K + s2 * 2 + q] = acc;
    }
}

// Per quasi-eigenstate q = e*K + s: <psi0_s|QE_q> (re, im) and |QE_q|^2.
// Grid (chunks, q tiles of 256); part[c][q][3].
__global__ void __launch_bounds__(256)
k_qe_stats(const int64_t* __restrict__ chunk, const double2* __restrict__ qe, int64_t nq, int K,
           const double2* __restrict__ psi0, int64_t n, const int* __restrict__ perm,
           double* __restrict__ part) {
    __shared__ double sh[256 * 3];
    const int c = remap_chunk(blockIdx.x, gridDim.x);
    const int64_t r0 = chunk[c] * 64, r1 = std::min<int64_t>(chunk[c + 1] * 64, n);
    const int64_t qb = (int64_t)blockIdx.y * 256;
    const int nqt = (int)std::min<int64_t>(256, nq - qb);
    const int ng = 256 / nqt;
    const int qi = threadIdx.x % nqt, grp = threadIdx.x / nqt;
    const int64_t q = qb + qi;
    const int s = (int)(q % K);
    double dr = 0, di = 0, nn = 0;
    if (grp < ng) {
        for (int64_t r = r0 + grp; r < r1; r += ng) {
            const double2 v = qe[r * nq + q];
            const double2 z = psi0[(int64_t)s * n + (perm ? (int64_t)perm[r] : r)];
            dr = fma(z.x, v.x, fma(z.y, v.y, dr));
            di = fma(z.x, v.y, fma(-z.y, v.x, di));
            nn = fma(v.x, v.x, fma(v.y, v.y, nn));
        }
    }
    sh[threadIdx.x * 3 + 0] = dr;
    sh[threadIdx.x * 3 + 1] = di;
    sh[threadIdx.x * 3 + 2] = nn;
    __syncthreads();
    // nqt * 3 sums (up to 768) over 256 threads
    for (int o = threadIdx.x; o < nqt * 3; o += blockDim.x) {
        const int q2 = o / 3, k = o % 3;
        double acc = 0.0;
        for (int g = 0; g < ng; ++g) acc += sh[(g * nqt + q2) * 3 + k];
        part[((int64_t)c * nq + qb + q2) * 3 + k] = acc;
    }
}

// Sum partial rows part[c][nv] over all chunks -> out[nv] (one thread per value).
__global__ void k_sum_rows(const double* __restrict__ part, int nchunks, int64_t nv,
                           double* __restrict__ out) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nv) return;
    double acc = 0.0;
    for (int c = 0; c < nchunks; ++c) acc += part[(int64_t)c * nv + q];
    out[q] = acc;
}

// |QE(r, e, s)|^2 in the caller's row order: out[(s * n_en + e) * n + perm[r]]
// (tbpm.f90:820-825, one sample per state s).
__global__ void k_qe_abs2(const double2* __restrict__ qe, int64_t n, int K, int n_en,
                          const int* __restrict__ perm, double* __restrict__ out) {
    const int64_t nq = (int64_t)n_en * K;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n * nq) return;
    const int64_t r = i / nq;
    const int q = (int)(i - r * nq);
    const int e = q / K, s = q - e * K;
    const double2 v = qe[i];
    const int64_t c = perm ? (int64_t)perm[r] : r;
    out[((int64_t)s * n_en + e) * n + c] = v.x * v.x + v.y * v.y;
}

__global__ void k_qe_scale(double2* __restrict__ qe, int64_t total, int64_t nq,
                           const double* __restrict__ div) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    const double d = div[i % nq];
    double2 v = qe[i];
    v.x = v.x / d;
    v.y = v.y / d;
    qe[i] = v;
}

// Y[r][a*nq + q] = sum_j J_a[r,j] (QE[col][q] / adiv[q])  (tbpm.f90:680-684)
template <typename R>
__global__ void k_dc_y(const int* __restrict__ cols, const typename CplxT<R>::T* __restrict__ jx,
                       const typename CplxT<R>::T* __restrict__ jy,
                       const int64_t* __restrict__ soff, const int* __restrict__ sw, int64_t n_pad,
                       const double2* __restrict__ qe, int64_t nq, const double* __restrict__ adiv,
                       double2* __restrict__ Y) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_pad * 2 * nq) return;
    const int64_t r = i / (2 * nq);
    const int64_t q2 = i - r * 2 * nq;
    const int a = q2 >= nq ? 1 : 0;
    const int64_t q = q2 - a * nq;
    const typename CplxT<R>::T* J = a ? jy : jx;
    const int64_t sl = r >> 6;
    const int lr = (int)(r & 63);
    const int64_t off = soff[sl];
    const int w = sw[sl];
    const double dv = adiv[q];
    double yr = 0.0, yi = 0.0;
    for (int j = 0; j < w; ++j) {
        const int64_t e = off + (int64_t)j * 64 + lr;
        const int col = cols[e];
        const auto h = J[e];
        const double2 v = qe[(int64_t)col * nq + q];
        const double vr = v.x / dv, vi = v.y / dv;
        const double hr = h.x, hi = h.y;
        yr = hr * vr - hi * vi + yr;
        yi = hr * vi + hi * vr + yi;
    }
    Y[i] = make_double2(yr, yi);
}

// dc[a][q] = <X2[a*K + s] | Y[a*nq + q]>, q = e*K + s.  Grid (chunks, tiles).
template <typename R>
__global__ void __launch_bounds__(256)
k_dcdot(const int64_t* __restrict__ chunk, const typename CplxT<R>::T* __restrict__ X, int Kp,
        int K, const double2* __restrict__ Y, int64_t nq2, int64_t nq,
        double* __restrict__ part) {
    __shared__ double sh[256 * 2];
    const int c = remap_chunk(blockIdx.x, gridDim.x);
    const int64_t r0 = chunk[c] * 64, r1 = chunk[c + 1] * 64;
    const int64_t qb = (int64_t)blockIdx.y * 256;
    const int nqt = (int)std::min<int64_t>(256, nq2 - qb);
    const int ng = 256 / nqt;
    const int qi = threadIdx.x % nqt, grp = threadIdx.x / nqt;
    const int64_t q2 = qb + qi;
    const int a = q2 >= nq ? 1 : 0;
    const int s = (int)((q2 - a * nq) % K);
    const int slot = a * K + s;
    double dr = 0, di = 0;
    if (grp < ng) {
        for (int64_t r = r0 + grp; r < r1; r += ng) {
            const auto x = X[r * Kp + slot];
            const double2 y = Y[r * nq2 + q2];
            const double xr = x.x, xi = x.y;
            dr = fma(xr, y.x, fma(xi, y.y, dr));
            di = fma(xr, y.y, fma(-xi, y.x, di));
        }
    }
    sh[threadIdx.x * 2 + 0] = dr;
    sh[threadIdx.x * 2 + 1] = di;
    __syncthreads();
    // nqt * 2 sums (up to 512) over 256 threads
    for (int o = threadIdx.x; o < nqt * 2; o += blockDim.x) {
        const int qq = o / 2, k = o % 2;
        double acc = 0.0;
        for (int g = 0; g < ng; ++g) acc += sh[(g * nqt + qq) * 2 + k];
        part[((int64_t)c * nq2 + qb + qq) * 2 + k] = acc;
    }
}

// Dynamic polarisation, per sample s: slots s (psi1), K+s (psi2):
// part[c][s] = Im sum_r conj(psi2_r) rho_r psi1_r, rho_r = e^{i q.r_r}
// (inner_prod(psi2, density_q .pMul. psi1), tbpm.f90:455-478).
template <typename R>
__global__ void __launch_bounds__(256)
k_rhocorr(const int64_t* __restrict__ chunk, const typename CplxT<R>::T* __restrict__ X, int Kp,
          int K, const double2* __restrict__ rho, double* __restrict__ part) {
    __shared__ double sh[256];
    const int c = remap_chunk(blockIdx.x, gridDim.x);
    const int64_t r0 = chunk[c] * 64, r1 = chunk[c + 1] * 64;
    const int ng = 256 / K;
    const int s = threadIdx.x % K, grp = threadIdx.x / K;
    double d = 0.0;
    if (grp < ng) {
        for (int64_t r = r0 + grp; r < r1; r += ng) {
            const auto a = X[r * Kp + s];
            const auto b = X[r * Kp + K + s];
            const double2 q = rho[r];
            const double ar = a.x, ai = a.y;
            const double wr = q.x * ar - q.y * ai, wi = q.x * ai + q.y * ar;
            d = fma((double)b.x, wi, fma(-(double)b.y, wr, d));
        }
    }
    sh[threadIdx.x] = d;
    __syncthreads();
    if (threadIdx.x < K) {
        double acc = 0.0;
        for (int g = 0; g < ng; ++g) acc += sh[g * K + threadIdx.x];
        part[(int64_t)c * K + threadIdx.x] = acc;
    }
}

// Slots [0, K) of X times conj(rho) per layout row (density(-q) .pMul. wf0).
template <typename R>
__global__ void k_rho_conj_mul(typename CplxT<R>::T* __restrict__ X, int Kp, int K,
                               const double2* __restrict__ rho, int64_t n_pad) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_pad * K) return;
    const int64_t r = i / K;
    const int s = (int)(i - r * K);
    auto v = X[r * Kp + s];
    const double2 q = rho[r];
    const double vr = v.x, vi = v.y;
    v.x = (R)(q.x * vr + q.y * vi);
    v.y = (R)(q.x * vi - q.y * vr);
    X[r * Kp + s] = v;
}

// ---------------------------------------------------------------------------
// Drivers
// ---------------------------------------------------------------------------
static bool broken(double n, double ref, double thr) {
    const double d = std::fabs(n - ref);
    return std::isnan(d) || !std::isfinite(d) || d > thr;
}

static int chunks_for(tbg_handle* h, int64_t nslices) { return chunk_target(h, nslices); }

template <typename R>
void corr_ac_impl(tbg_handle* h, const tbg_ham* H, const double* bes, int nb, const double* cF,
                  int nF, const double* cO, int nO, const double* psi0, int K, int T,
                  int chk_step, double chk_thr, double* corr, int* fail_step) {
    using C = typename CplxT<R>::T;
    Engine<R> e;
    e.init(h, view_of(H), {0, H->nslices}, 3 * K, chunks_for(h, H->nslices), H);
    const int Kp = e.Kp;
    const hipStream_t st = h->stream;
    DevBuf psi(e.vbytes);
    launch_fill_zero(psi.p, e.vbytes, st);
    launch_pack<R>(psi0, K, H->n, psi.p, Kp, 0, H->perm_ptr(), st);
    // psi1_x = J_x psi0, psi1_y = J_y psi0, psi2 = psi0 (tbpm.f90:279-281)
    launch_spmv<R>(H, H->jx.p, psi.p, Kp, 0, e.X.p, Kp, 0, K, st);
    launch_spmv<R>(H, H->jy.p, psi.p, Kp, 0, e.X.p, Kp, K, K, st);
    launch_pack<R>(psi0, K, H->n, e.X.p, Kp, 2 * K, H->perm_ptr(), st);
    // Fermi operators: class 0 = 1 - F on psi1, class 1 = F on psi2.
    uint64_t cls = 0;
    for (int k = 2 * K; k < 3 * K; ++k) cls |= 1ull << k;
    const StepPoly fp = cheb_poly(cO, nO, cF, nF);
    {
        TimingHold hold(h);
        e.step(fp, 0, cls, 0, nullptr, nullptr);
    }
    // norm_ref of psi2 (tbpm.f90:306) and per-step norms of every slot.
    const int nvn = 3 * Kp;
    DevBuf nbuf((size_t)T * nvn * sizeof(double));
    DevBuf abuf((size_t)T * 8 * K * sizeof(double));
    TBG_HIP(hipMemsetAsync(abuf.p, 0, (size_t)T * 8 * K * sizeof(double), st));
    DevBuf counter(sizeof(int));
    TBG_HIP(hipMemsetAsync(counter.p, 0, sizeof(int), st));
    launch_dotnorm<R>(e.plan, e.X.p, e.X.p, Kp, Kp, e.part.template as<double>(), st);
    launch_reduce(e.part.template as<double>(), nvn, e.plan.inst_chunk_dev.template as<int>(), 1,
                  nbuf.as<double>(), 0, counter.as<int>(), 0, st);
    const StepPoly sp = time_poly(bes, nb);
    auto after = [&]() {
        launch_reduce(e.part.template as<double>(), nvn, e.epi_range(), 1, nbuf.as<double>(), 0,
                      counter.as<int>(), 0, st);
        timed_launch(h, kTimeAcCorr, [&] {
            hipLaunchKernelGGL(k_accorr<R>, dim3(e.plan.nchunks), dim3(256), 0, st,
                               H->cols.as<int>(), H->jx.as<C>(), H->jy.as<C>(),
                               H->slice_off.as<int64_t>(), H->slice_w.as<int>(),
                               e.plan.chunk.template as<int64_t>(), e.X.template as<C>(), Kp, K,
                               e.part.template as<double>());
        });
        launch_reduce(e.part.template as<double>(), 8 * K, e.plan.inst_chunk_dev.template as<int>(),
                      1, abuf.as<double>(), 0, counter.as<int>(), 0, st);
    };
    // t = 2..T (tbpm.f90:308-345); t = 1 is never stored.
    if (T >= 2) e.run(T - 1, sp, 0, 0, F_NORM, nullptr, counter.as<int>(), after);
    std::vector<double> ha((size_t)T * 8 * K), hn((size_t)T * nvn);
    TBG_HIP(hipMemcpyAsync(ha.data(), abuf.p, ha.size() * sizeof(double), hipMemcpyDeviceToHost, st));
    TBG_HIP(hipMemcpyAsync(hn.data(), nbuf.p, hn.size() * sizeof(double), hipMemcpyDeviceToHost, st));
    TBG_HIP(hipStreamSynchronize(st));
    std::vector<double> out((size_t)K * 4 * T * 2);
    for (int s = 0; s < K; ++s)
        for (int q = 0; q < 4; ++q)
            for (int t = 0; t < T; ++t)
                for (int ri = 0; ri < 2; ++ri)
                    out[(((size_t)s * 4 + q) * T + t) * 2 + ri] = ha[(size_t)t * 8 * K + s * 8 + 2 * q + ri];
    TBG_HIP(hipMemcpyAsync(corr, out.data(), out.size() * sizeof(double), hipMemcpyHostToDevice, st));
    TBG_HIP(hipStreamSynchronize(st));
    // Check at loop index chk_step on psi2 after chk_step-2 steps.
    if (chk_step >= 2 && chk_step <= T) {
        for (int s = 0; s < K; ++s) {
            const int slot = 2 * K + s;
            const double ref = hn[slot * 3 + 2];
            const double nn = hn[(size_t)(chk_step - 2) * nvn + slot * 3 + 2];
            if (broken(nn, ref, chk_thr)) {
                if (fail_step) *fail_step = chk_step;
                throw Error{TBG_ENORM, "wave function norm broken at step " +
                                           std::to_string(chk_step) + " (state " +
                                           std::to_string(s) + "); check the Hermiticity of the "
                                           "Hamiltonian or increase the rescale factor"};
            }
        }
    }
}

// tbpm_dyn_pol (tbpm.f90:355-493) on caller states: per q-point, psi1 =
// (1 - F)(H) (density(-q) .pMul. psi0) and psi2 = F(H) psi0, propagated
// together; corr(s, q, t) = Im <psi2 | density(q) .pMul. psi1> at t = 1
// (before any step) .. T.  Norm check at loop index chk_step (2..T) on
// psi2 after chk_step-2 steps against its norm after the Fermi operator.
template <typename R>
void dyn_pol_impl(tbg_handle* h, const tbg_ham* H, const double* bes, int nb, const double* cF,
                  int nF, const double* cO, int nO, const double* psi0, int K, int T,
                  const double* q_points, int n_q, const double* pos, int chk_step,
                  double chk_thr, double* corr, int* fail_step) {
    using C = typename CplxT<R>::T;
    const hipStream_t st = h->stream;
    Engine<R> e;
    e.init(h, view_of(H), {0, H->nslices}, 2 * K, chunks_for(h, H->nslices), H);
    const int Kp = e.Kp;
    const int64_t n = H->n, n_pad = e.n_pad;
    uint64_t cls = 0;
    for (int k = K; k < 2 * K; ++k) cls |= 1ull << k;
    const StepPoly fp = cheb_poly(cO, nO, cF, nF);
    const StepPoly sp = time_poly(bes, nb);
    const int nvn = 3 * Kp;
    DevBuf rho((size_t)n_pad * sizeof(double2)), rho_in((size_t)n * sizeof(double2));
    DevBuf nbuf((size_t)T * nvn * sizeof(double));
    DevBuf cbuf((size_t)T * K * sizeof(double));
    DevBuf counter(sizeof(int));
    std::vector<double> hr((size_t)n * 2), hc((size_t)T * K), hn((size_t)T * nvn);
    std::vector<double> out((size_t)K * n_q * T);
    auto corr_now = [&]() {
        hipLaunchKernelGGL(k_rhocorr<R>, dim3(e.plan.nchunks), dim3(256), 0, st,
                           e.plan.chunk.template as<int64_t>(), e.X.template as<C>(), Kp, K,
                           rho.as<double2>(), e.part.template as<double>());
        launch_reduce(e.part.template as<double>(), K, e.plan.inst_chunk_dev.template as<int>(),
                      1, cbuf.as<double>(), 0, counter.as<int>(), 0, st);
    };
    auto after = [&]() {
        launch_reduce(e.part.template as<double>(), nvn, e.epi_range(), 1, nbuf.as<double>(), 0,
                      counter.as<int>(), 0, st);
        corr_now();
    };
    for (int iq = 0; iq < n_q; ++iq) {
        // density(q) = e^{i q.r} per orbital (funcs.f90:125-149), caller order
        const double* q = q_points + (size_t)iq * 3;
        for (int64_t j = 0; j < n; ++j) {
            const double pw = q[0] * pos[j * 3 + 0] + q[1] * pos[j * 3 + 1] + q[2] * pos[j * 3 + 2];
            hr[j * 2 + 0] = std::cos(pw);
            hr[j * 2 + 1] = std::sin(pw);
        }
        TBG_HIP(hipMemcpyAsync(rho_in.p, hr.data(), hr.size() * sizeof(double),
                               hipMemcpyHostToDevice, st));
        launch_fill_zero(rho.p, rho.bytes, st);
        launch_pack<double>(rho_in.as<double>(), 1, n, rho.p, 1, 0, H->perm_ptr(), st);
        // psi1 = (1 - F) density(-q) psi0 in slots [0, K), psi2 = F psi0 in [K, 2K)
        launch_fill_zero(e.X.p, e.vbytes, st);
        launch_pack<R>(psi0, K, n, e.X.p, Kp, 0, H->perm_ptr(), st);
        launch_pack<R>(psi0, K, n, e.X.p, Kp, K, H->perm_ptr(), st);
        hipLaunchKernelGGL(k_rho_conj_mul<R>, dim3(nblk(n_pad * K)), dim3(256), 0, st,
                           e.X.template as<C>(), Kp, K, rho.as<double2>(), n_pad);
        TBG_HIP(hipGetLastError());
        {
            TimingHold hold(h);
            e.step(fp, 0, cls, 0, nullptr, nullptr);
        }
        TBG_HIP(hipMemsetAsync(counter.p, 0, sizeof(int), st));
        launch_dotnorm<R>(e.plan, e.X.p, e.X.p, Kp, Kp, e.part.template as<double>(), st);
        launch_reduce(e.part.template as<double>(), nvn, e.plan.inst_chunk_dev.template as<int>(),
                      1, nbuf.as<double>(), 0, counter.as<int>(), 0, st);
        corr_now();   // t = 1 (tbpm.f90:459-466)
        if (T >= 2) e.run(T - 1, sp, 0, 0, F_NORM, nullptr, counter.as<int>(), after);
        TBG_HIP(hipMemcpyAsync(hc.data(), cbuf.p, hc.size() * sizeof(double),
                               hipMemcpyDeviceToHost, st));
        TBG_HIP(hipMemcpyAsync(hn.data(), nbuf.p, hn.size() * sizeof(double),
                               hipMemcpyDeviceToHost, st));
        TBG_HIP(hipStreamSynchronize(st));
        for (int s = 0; s < K; ++s)
            for (int t = 0; t < T; ++t) out[((size_t)s * n_q + iq) * T + t] = hc[(size_t)t * K + s];
        if (chk_step >= 2 && chk_step <= T) {
            for (int s = 0; s < K; ++s) {
                const int slot = K + s;
                const double ref = hn[slot * 3 + 2];
                const double nn = hn[(size_t)(chk_step - 2) * nvn + slot * 3 + 2];
                if (broken(nn, ref, chk_thr)) {
                    if (fail_step) *fail_step = chk_step;
                    throw Error{TBG_ENORM, "wave function norm broken at step " +
                                               std::to_string(chk_step) + " (state " +
                                               std::to_string(s) + ", q-point " +
                                               std::to_string(iq) + "); check the Hermiticity "
                                               "of the Hamiltonian or increase the rescale factor"};
                }
            }
        }
    }
    TBG_HIP(hipMemcpyAsync(corr, out.data(), out.size() * sizeof(double), hipMemcpyHostToDevice,
                           st));
    TBG_HIP(hipStreamSynchronize(st));
}

template <typename R>
void corr_dc_impl(tbg_handle* h, const tbg_ham* H, const double* bes, int nb, const double* psi0,
                  int K, int T, double t_step, const double* energies, const int64_t* en_inds,
                  int n_en, int chk_step, double chk_thr, double* dos_corr, double* dc_corr,
                  int* fail_step, double* qe_abs2) {
    using C = typename CplxT<R>::T;
    const hipStream_t st = h->stream;
    Engine<R> e;
    e.init(h, view_of(H), {0, H->nslices}, 2 * K, chunks_for(h, H->nslices), H);
    const int Kp = e.Kp;
    const int64_t n = H->n, n_pad = e.n_pad;
    const int64_t nq = (int64_t)n_en * K;
    DevBuf psi(e.vbytes);
    launch_fill_zero(psi.p, e.vbytes, st);
    launch_pack<R>(psi0, K, n, psi.p, Kp, 0, H->perm_ptr(), st);
    launch_pack<R>(psi0, K, n, e.X.p, Kp, 0, H->perm_ptr(), st);
    launch_pack<R>(psi0, K, n, e.X.p, Kp, K, H->perm_ptr(), st);
    uint64_t bwd = 0;
    for (int k = K; k < 2 * K; ++k) bwd |= 1ull << k;
    DevBuf qe((size_t)std::max<int64_t>(n_pad * nq, 1) * sizeof(double2));
    if (nq > 0)
        hipLaunchKernelGGL(k_qe_init, dim3(nblk(n_pad * nq)), dim3(256), 0, st,
                           reinterpret_cast<const double2*>(psi0), n, K, n_en, n_pad,
                           H->perm_ptr(), qe.as<double2>());
    // Phase tables P(t, e) = exp(i E t dt), W(t) = Hanning (tbpm.f90:625-629).
    std::vector<double> ptab((size_t)T * std::max(n_en, 1) * 2), wtab(T);
    const double pi = 3.14159265358979323846;
    for (int t = 1; t <= T; ++t) {
        wtab[t - 1] = 0.5 * (1 + std::cos(pi * t / T));
        for (int i = 0; i < n_en; ++i) {
            const double th = energies[en_inds[i]] * t * t_step;
            ptab[((size_t)(t - 1) * n_en + i) * 2 + 0] = std::cos(th);
            ptab[((size_t)(t - 1) * n_en + i) * 2 + 1] = std::sin(th);
        }
    }
    DevBuf dptab(ptab.size() * sizeof(double)), dwtab(wtab.size() * sizeof(double));
    TBG_HIP(hipMemcpyAsync(dptab.p, ptab.data(), ptab.size() * sizeof(double), hipMemcpyHostToDevice, st));
    TBG_HIP(hipMemcpyAsync(dwtab.p, wtab.data(), wtab.size() * sizeof(double), hipMemcpyHostToDevice, st));
    const int nvn = 3 * Kp;
    DevBuf nbuf((size_t)(T + 1) * nvn * sizeof(double));
    DevBuf dbuf((size_t)T * 2 * K * sizeof(double));
    DevBuf counter(sizeof(int));
    TBG_HIP(hipMemsetAsync(counter.p, 0, sizeof(int), st));
    launch_dotnorm<R>(e.plan, e.X.p, e.X.p, Kp, Kp, e.part.template as<double>(), st);
    launch_reduce(e.part.template as<double>(), nvn, e.plan.inst_chunk_dev.template as<int>(), 1,
                  nbuf.as<double>(), 0, counter.as<int>(), 0, st);
    const StepPoly sp = time_poly(bes, nb);
    // Check at t == chk_step on psi(+-) after chk_step-1 steps vs <psi0|psi0>
    // (tbpm.f90:608-611, 772-775).
    auto check_norms = [&]() {
        if (chk_step < 1 || chk_step > T) return;
        std::vector<double> hn((size_t)(T + 1) * nvn);
        TBG_HIP(hipMemcpyAsync(hn.data(), nbuf.p, hn.size() * sizeof(double),
                               hipMemcpyDeviceToHost, st));
        TBG_HIP(hipStreamSynchronize(st));
        for (int k = 0; k < 2 * K; ++k) {
            const double ref = hn[k * 3 + 2];
            const double nn = hn[(size_t)(chk_step - 1) * nvn + k * 3 + 2];
            if (broken(nn, ref, chk_thr)) {
                if (fail_step) *fail_step = chk_step;
                throw Error{TBG_ENORM, "wave function norm broken at step " +
                                           std::to_string(chk_step) +
                                           "; check the Hermiticity of the Hamiltonian or "
                                           "increase the rescale factor"};
            }
        }
    };
    auto after1 = [&]() {
        launch_reduce(e.part.template as<double>(), nvn, e.epi_range(), 1, nbuf.as<double>(), 0,
                      counter.as<int>(), 0, st);
        timed_launch(h, kTimeQe, [&] {
            hipLaunchKernelGGL(k_qe<R>, dim3(e.plan.nchunks), dim3(256), 0, st,
                               e.plan.chunk.template as<int64_t>(), e.X.template as<C>(), Kp, K,
                               psi.as<C>(), qe.as<double2>(), n_en, dptab.as<double2>(),
                               dwtab.as<double>(), counter.as<int>(),
                               e.part.template as<double>());
        });
        launch_reduce(e.part.template as<double>(), 2 * K, e.plan.inst_chunk_dev.template as<int>(),
                      1, dbuf.as<double>(), 0, counter.as<int>(), -1, st);
    };
    e.run(T, sp, bwd, 0, F_NORM, nullptr, counter.as<int>(), after1);
    // Normalise the quasi-eigenstates (tbpm.f90:642-647) and form
    // a = |<psi0|QE>| on the normalised states (tbpm.f90:667).
    std::vector<double> adiv(std::max<int64_t>(nq, 1));
    DevBuf dadiv(adiv.size() * sizeof(double));
    if (nq > 0) {
        const int64_t ntiles = (nq + 255) / 256;
        DevBuf qpart((size_t)e.plan.nchunks * nq * 3 * sizeof(double));
        DevBuf qsum((size_t)nq * 3 * sizeof(double));
        std::vector<double> hs((size_t)nq * 3);
        for (int pass = 0; pass < 2; ++pass) {
            hipLaunchKernelGGL(k_qe_stats, dim3(e.plan.nchunks, (unsigned)ntiles), dim3(256), 0, st,
                               e.plan.chunk.template as<int64_t>(), qe.as<double2>(), nq, K,
                               reinterpret_cast<const double2*>(psi0), n, H->perm_ptr(),
                               qpart.as<double>());
            TBG_HIP(hipGetLastError());
            hipLaunchKernelGGL(k_sum_rows, dim3(nblk(nq * 3)), dim3(256), 0, st, qpart.as<double>(),
                               e.plan.nchunks, nq * 3, qsum.as<double>());
            TBG_HIP(hipGetLastError());
            TBG_HIP(hipMemcpyAsync(hs.data(), qsum.p, hs.size() * sizeof(double),
                                   hipMemcpyDeviceToHost, st));
            TBG_HIP(hipStreamSynchronize(st));
            if (pass == 0) {
                for (int64_t q = 0; q < nq; ++q) adiv[q] = std::sqrt(hs[q * 3 + 2]);
                TBG_HIP(hipMemcpyAsync(dadiv.p, adiv.data(), nq * sizeof(double),
                                       hipMemcpyHostToDevice, st));
                hipLaunchKernelGGL(k_qe_scale, dim3(nblk(n_pad * nq)), dim3(256), 0, st,
                                   qe.as<double2>(), n_pad * nq, nq, dadiv.as<double>());
                TBG_HIP(hipGetLastError());
                if (qe_abs2 != nullptr) {   // quasi-eigenstates: done after normalising
                    hipLaunchKernelGGL(k_qe_abs2, dim3(nblk(n * nq)), dim3(256), 0, st,
                                       qe.as<double2>(), n, K, n_en, H->perm_ptr(), qe_abs2);
                    TBG_HIP(hipGetLastError());
                    TBG_HIP(hipStreamSynchronize(st));
                    check_norms();
                    return;
                }
            } else {
                for (int64_t q = 0; q < nq; ++q) adiv[q] = std::hypot(hs[q * 3 + 0], hs[q * 3 + 1]);
                TBG_HIP(hipMemcpyAsync(dadiv.p, adiv.data(), nq * sizeof(double),
                                       hipMemcpyHostToDevice, st));
            }
        }
    }
    // Y_a = J_a (QE / a), time independent.
    DevBuf Y((size_t)std::max<int64_t>(n_pad * 2 * nq, 1) * sizeof(double2));
    if (nq > 0) {
        hipLaunchKernelGGL(k_dc_y<R>, dim3(nblk(n_pad * 2 * nq)), dim3(256), 0, st,
                           H->cols.as<int>(), H->jx.as<C>(), H->jy.as<C>(),
                           H->slice_off.as<int64_t>(), H->slice_w.as<int>(), n_pad,
                           qe.as<double2>(), nq, dadiv.as<double>(), Y.as<double2>());
        TBG_HIP(hipGetLastError());
    }
    qe.release();
    // Phase 2: J_x psi0, J_y psi0 propagated forward (tbpm.f90:655-697).
    e.drop_graph();
    launch_fill_zero(e.X.p, e.vbytes, st);
    launch_spmv<R>(H, H->jx.p, psi.p, Kp, 0, e.X.p, Kp, 0, K, st);
    launch_spmv<R>(H, H->jy.p, psi.p, Kp, 0, e.X.p, Kp, K, K, st);
    const int64_t nq2 = 2 * nq;
    DevBuf cbuf((size_t)std::max<int64_t>(T * nq2 * 2, 1) * sizeof(double));
    DevBuf cpart((size_t)std::max<int64_t>((int64_t)e.plan.nchunks * nq2 * 2, 1) * sizeof(double));
    TBG_HIP(hipMemsetAsync(counter.p, 0, sizeof(int), st));
    const int64_t ntiles2 = (nq2 + 255) / 256;
    auto after2 = [&]() {
        if (nq2 == 0) return;
        timed_launch(h, kTimeDcDot, [&] {
            hipLaunchKernelGGL(k_dcdot<R>, dim3(e.plan.nchunks, (unsigned)ntiles2), dim3(256), 0,
                               st, e.plan.chunk.template as<int64_t>(), e.X.template as<C>(), Kp,
                               K, Y.as<double2>(), nq2, nq, cpart.as<double>());
        });
        launch_reduce(cpart.as<double>(), (int)(nq2 * 2), e.plan.inst_chunk_dev.template as<int>(),
                      1, cbuf.as<double>(), 0, counter.as<int>(), 0, st);
    };
    after2();  // t = 1, unpropagated
    if (T >= 2) e.run(T - 1, sp, 0, 0, 0, nullptr, counter.as<int>(), after2);
    std::vector<double> hd((size_t)T * 2 * K), hc((size_t)T * nq2 * 2);
    TBG_HIP(hipMemcpyAsync(hd.data(), dbuf.p, hd.size() * sizeof(double), hipMemcpyDeviceToHost, st));
    if (nq2 > 0)
        TBG_HIP(hipMemcpyAsync(hc.data(), cbuf.p, hc.size() * sizeof(double), hipMemcpyDeviceToHost, st));
    TBG_HIP(hipStreamSynchronize(st));
    std::vector<double> od((size_t)K * T * 2), oc((size_t)K * 2 * std::max(n_en, 0) * T * 2);
    for (int s = 0; s < K; ++s)
        for (int t = 0; t < T; ++t)
            for (int ri = 0; ri < 2; ++ri)
                od[((size_t)s * T + t) * 2 + ri] = hd[(size_t)t * 2 * K + s * 2 + ri];
    for (int s = 0; s < K; ++s)
        for (int a = 0; a < 2; ++a)
            for (int i = 0; i < n_en; ++i)
                for (int t = 0; t < T; ++t)
                    for (int ri = 0; ri < 2; ++ri)
                        oc[((((size_t)s * 2 + a) * n_en + i) * T + t) * 2 + ri] =
                            hc[(size_t)t * nq2 * 2 + (a * nq + (int64_t)i * K + s) * 2 + ri];
    TBG_HIP(hipMemcpyAsync(dos_corr, od.data(), od.size() * sizeof(double), hipMemcpyHostToDevice, st));
    if (!oc.empty())
        TBG_HIP(hipMemcpyAsync(dc_corr, oc.data(), oc.size() * sizeof(double), hipMemcpyHostToDevice, st));
    TBG_HIP(hipStreamSynchronize(st));
    check_norms();
}

template void corr_ac_impl<float>(tbg_handle*, const tbg_ham*, const double*, int, const double*,
                                  int, const double*, int, const double*, int, int, int, double,
                                  double*, int*);
template void corr_ac_impl<double>(tbg_handle*, const tbg_ham*, const double*, int, const double*,
                                   int, const double*, int, const double*, int, int, int, double,
                                   double*, int*);
template void dyn_pol_impl<float>(tbg_handle*, const tbg_ham*, const double*, int, const double*,
                                  int, const double*, int, const double*, int, int, const double*,
                                  int, const double*, int, double, double*, int*);
template void dyn_pol_impl<double>(tbg_handle*, const tbg_ham*, const double*, int,
                                   const double*, int, const double*, int, const double*, int, int,
                                   const double*, int, const double*, int, double, double*, int*);
template void corr_dc_impl<float>(tbg_handle*, const tbg_ham*, const double*, int, const double*,
                                  int, int, double, const double*, const int64_t*, int, int, double,
                                  double*, double*, int*, double*);
template void corr_dc_impl<double>(tbg_handle*, const tbg_ham*, const double*, int, const double*,
                                   int, int, double, const double*, const int64_t*, int, int,
                                   double, double*, double*, int*, double*);

}  // namespace tbg
