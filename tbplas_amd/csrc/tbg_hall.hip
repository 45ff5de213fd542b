// Kubo-Bastin Hall conductivity moments (KPM): mu_mn of tbpm_kbdc
// (tbplas/fortran/tbpm.f90:836-960), and the Haydock recursion
// coefficients of ldos_haydock (tbpm.f90:964-1024, Haydock_coef
// propagation.f90:85-155), on the device.
//
// For each state psi: W_j = J_b T_j(H~) psi (j < N, b = x or y), P_i =
// T_i(H~) J_x psi (i < N) and mu(i, j) = g_i g_j <P_i | W_j> with the Jackson
// weights g (kpm.f90:8-29).  The Chebyshev vectors come from the term
// kernel (T_j = 2 H~ T_{j-1} - T_{j-2}, in place over T_{j-2}); the
// currents from the slice-ELL SpMV; the N x N Gram matrix P^H W is a plain
// complex128 GEMM (rocBLAS zgemm), issued per block of 256 rows of mu while
// the next block of P is formed.  Complex128 only: the moments are sums of
// N^2 products that complex64 would leave at ~1e-7 of their size.
#include <rocblas/rocblas.h>

#include <cmath>
#include <vector>

#include "tbg_engine.h"

namespace tbg {

namespace {

// mu[i + j N] *= g_i g_j  (the reference's g(:) * g(j) * corr_mu(:, j))
__global__ void k_jackson_scale(double2* __restrict__ mu, int N, const double* __restrict__ g) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)N * N) return;
    const int i = (int)(t % N), j = (int)(t / N);
    const double w = g[i] * g[j];
    double2 v = mu[t];
    v.x = w * v.x;
    v.y = w * v.y;
    mu[t] = v;
}

struct Blas {
    rocblas_handle b = nullptr;
    explicit Blas(hipStream_t s) {
        if (rocblas_create_handle(&b) != rocblas_status_success)
            throw Error{TBG_EHIP, "rocblas_create_handle failed"};
        if (rocblas_set_stream(b, s) != rocblas_status_success) {
            rocblas_destroy_handle(b);
            throw Error{TBG_EHIP, "rocblas_set_stream failed"};
        }
    }
    ~Blas() {
        if (b) rocblas_destroy_handle(b);
    }
};

// Haydock step i on one vector (Kp = 1): buf -= a_i cur + b_{i-1} prev with
// a_i = dots[3i] + i dots[3i+1] and b_{i-1} = sqrt(nrm[3(i-1)+2]) read on
// the device (axpbypz, propagation.f90:144).
__global__ void k_lz_update(double2* __restrict__ buf, const double2* __restrict__ cur,
                            const double2* __restrict__ prev, int64_t np,
                            const double* __restrict__ dots, const double* __restrict__ nrm,
                            int i) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= np) return;
    const double ar = dots[3 * i], ai = dots[3 * i + 1];
    const double b = i > 0 ? sqrt(nrm[3 * (i - 1) + 2]) : 0.0;
    const double2 c = cur[r], p = prev[r];
    double2 v = buf[r];
    v.x = v.x - (ar * c.x - ai * c.y) - b * p.x;
    v.y = v.y - (ar * c.y + ai * c.x) - b * p.y;
    buf[r] = v;
}

// buf /= b_i = sqrt(nrm[3i+2]) (propagation.f90:146)
__global__ void k_lz_scale(double2* __restrict__ buf, int64_t np, const double* __restrict__ nrm,
                           int i) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= np) return;
    const double b = sqrt(nrm[3 * i + 2]);
    double2 v = buf[r];
    v.x = v.x / b;
    v.y = v.y / b;
    buf[r] = v;
}

}  // namespace

void haydock_impl(tbg_handle* h, const tbg_ham* H, const double* psi0, int K, int depth,
                  double rescale, double* coefa, double* coefb) {
    const hipStream_t st = h->stream;
    Engine<double> e;
    e.init(h, view_of(H), {0, H->nslices}, 1, chunk_target(h, H->nslices), nullptr);
    TBG_REQUIRE(e.Kp == 1, "internal: one state per Haydock launch");
    const int64_t n = H->n, np = e.n_pad;
    const size_t vb = (size_t)np * sizeof(double2);
    DevBuf v0(vb), v1(vb), v2(vb);
    DevBuf dots((size_t)depth * 3 * sizeof(double)), nrm((size_t)depth * 3 * sizeof(double));
    double* part = e.part.template as<double>();
    const int* range = e.plan.inst_chunk_dev.template as<int>();
    const unsigned nb = (unsigned)((np + 255) / 256);
    std::vector<double> hd((size_t)depth * 3), hn((size_t)depth * 3);
    for (int s = 0; s < K; ++s) {
        void* cur = v0.p;
        void* prev = v1.p;
        void* buf = v2.p;
        launch_fill_zero(cur, vb, st);
        launch_fill_zero(prev, vb, st);
        launch_pack<double>(psi0 + (size_t)s * n * 2, 1, n, cur, 1, 0, H->perm_ptr(), st);
        for (int i = 0; i < depth; ++i) {
            TermArgs a = e.base_args(0, 0);   // buf = rescale H~ cur  (amv)
            a.pin = cur;
            a.pout = buf;
            a.ar[0] = a.ar[1] = rescale;
            a.ai[0] = a.ai[1] = 0.0;
            launch_term<double>(F_STORE, a, e.plan.nchunks, st);
            launch_dotnorm<double>(e.plan, cur, buf, 1, 1, part, st);   // a_i = <cur|buf>
            launch_reduce(part, 3, range, 1, dots.as<double>(), 0, nullptr, i, st);
            hipLaunchKernelGGL(k_lz_update, dim3(nb), dim3(256), 0, st,
                               reinterpret_cast<double2*>(buf),
                               reinterpret_cast<const double2*>(cur),
                               reinterpret_cast<const double2*>(prev), np, dots.as<double>(),
                               nrm.as<double>(), i);
            TBG_HIP(hipGetLastError());
            launch_dotnorm<double>(e.plan, buf, buf, 1, 1, part, st);   // b_i = |buf|
            launch_reduce(part, 3, range, 1, nrm.as<double>(), 0, nullptr, i, st);
            hipLaunchKernelGGL(k_lz_scale, dim3(nb), dim3(256), 0, st,
                               reinterpret_cast<double2*>(buf), np, nrm.as<double>(), i);
            TBG_HIP(hipGetLastError());
            void* t = prev;
            prev = cur;
            cur = buf;
            buf = t;
        }
        TBG_HIP(hipMemcpyAsync(hd.data(), dots.p, hd.size() * sizeof(double),
                               hipMemcpyDeviceToHost, st));
        TBG_HIP(hipMemcpyAsync(hn.data(), nrm.p, hn.size() * sizeof(double),
                               hipMemcpyDeviceToHost, st));
        TBG_HIP(hipStreamSynchronize(st));
        for (int i = 0; i < depth; ++i) {
            coefa[((size_t)s * depth + i) * 2 + 0] = hd[(size_t)i * 3 + 0];
            coefa[((size_t)s * depth + i) * 2 + 1] = hd[(size_t)i * 3 + 1];
            coefb[(size_t)s * depth + i] = std::sqrt(hn[(size_t)i * 3 + 2]);
        }
    }
}

void hall_mu_impl(tbg_handle* h, const tbg_ham* H, const double* psi0, int K, int N, int dir,
                  double* mu) {
    const hipStream_t st = h->stream;
    Engine<double> e;
    e.init(h, view_of(H), {0, H->nslices}, 1, chunk_target(h, H->nslices), nullptr);
    TBG_REQUIRE(e.Kp == 1, "internal: one state per Hall launch");
    const int64_t n = H->n, np = e.n_pad;
    const size_t vb = (size_t)np * sizeof(double2);
    // Jackson kernel (kpm.f90:20-27)
    std::vector<double> g(N);
    g[0] = 0.5;
    const double q = M_PI / N, cotq = std::cos(q) / std::sin(q);
    for (int k = 1; k < N; ++k) g[k] = ((N - k) * std::cos(q * k) + cotq * std::sin(q * k)) / N;
    DevBuf dg(g.size() * sizeof(double));
    TBG_HIP(hipMemcpyAsync(dg.p, g.data(), g.size() * sizeof(double), hipMemcpyHostToDevice, st));
    const int B = std::min(N, 256);
    DevBuf W((size_t)N * vb), Pb((size_t)B * vb), T0(vb), T1(vb), psi(vb);
    const void* jb = dir == 1 ? H->jx.p : H->jy.p;
    auto col = [&](DevBuf& m, int j) { return reinterpret_cast<char*>(m.p) + (size_t)j * vb; };
    // T_{m} = alpha H~ pin + beta pprev, written to pout (pout may be pprev)
    auto term = [&](const void* pin, const void* pprev, void* pout, bool first) {
        TermArgs a = e.base_args(0, 0);
        a.pin = pin;
        a.pprev = pprev;
        a.pout = pout;
        a.ar[0] = a.ar[1] = first ? 1.0 : 2.0;
        a.ai[0] = a.ai[1] = 0.0;
        a.beta = -1.0;
        launch_term<double>(first ? F_STORE : (F_PREV | F_STORE), a, e.plan.nchunks, st);
    };
    Blas blas(st);
    const rocblas_double_complex one = {1.0, 0.0}, zero = {0.0, 0.0};
    for (int s = 0; s < K; ++s) {
        launch_fill_zero(psi.p, vb, st);
        launch_pack<double>(psi0 + (size_t)s * n * 2, 1, n, psi.p, 1, 0, H->perm_ptr(), st);
        // W_j = J_b T_j psi (tbpm.f90:892-922)
        void* cur = T0.p;   // T_{j-1}
        void* oth = T1.p;   // T_{j-2}; T_j is formed in place over it
        for (int j = 0; j < N; ++j) {
            const void* tj = psi.p;
            if (j == 1) {
                term(psi.p, nullptr, cur, true);
                tj = cur;
            } else if (j == 2) {
                term(cur, psi.p, oth, false);
                std::swap(cur, oth);
                tj = cur;
            } else if (j > 2) {
                term(cur, oth, oth, false);
                std::swap(cur, oth);
                tj = cur;
            }
            launch_spmv<double>(H, jb, tj, 1, 0, col(W, j), 1, 0, 1, st);
        }
        // P_i = T_i J_x psi in blocks of B columns, mu[i0.., :] = P^H W
        double* mus = mu + (size_t)s * N * N * 2;
        void* p0 = nullptr;   // P_{i-2}
        void* p1 = nullptr;   // P_{i-1}
        for (int i = 0; i < N; ++i) {
            void* pi = col(Pb, i % B);
            if (i == 0) {
                launch_spmv<double>(H, H->jx.p, psi.p, 1, 0, pi, 1, 0, 1, st);
            } else if (i == 1) {
                term(p1, nullptr, pi, true);
            } else {
                term(p1, p0, pi, false);
            }
            p0 = p1;
            p1 = pi;
            if (i % B == B - 1 || i == N - 1) {
                const int i0 = i - i % B, rows = i % B + 1;
                if (rocblas_zgemm(blas.b, rocblas_operation_conjugate_transpose,
                                  rocblas_operation_none, rows, N, (rocblas_int)np, &one,
                                  reinterpret_cast<const rocblas_double_complex*>(Pb.p),
                                  (rocblas_int)np,
                                  reinterpret_cast<const rocblas_double_complex*>(W.p),
                                  (rocblas_int)np, &zero,
                                  reinterpret_cast<rocblas_double_complex*>(mus) + i0, N) !=
                    rocblas_status_success)
                    throw Error{TBG_EHIP, "rocblas_zgemm failed"};
            }
        }
        const int64_t nn = (int64_t)N * N;
        hipLaunchKernelGGL(k_jackson_scale, dim3((unsigned)((nn + 255) / 256)), dim3(256), 0, st,
                           reinterpret_cast<double2*>(mus), N, dg.as<double>());
        TBG_HIP(hipGetLastError());
    }
    TBG_HIP(hipStreamSynchronize(st));
}

}  // namespace tbg
