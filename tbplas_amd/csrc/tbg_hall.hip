// Kubo-Bastin Hall conductivity moments (KPM): mu_mn of tbpm_kbdc
// (tbplas/fortran/tbpm.f90:836-960), and the Haydock recursion
// coefficients of ldos_haydock (tbpm.f90:964-1024, Haydock_coef
// propagation.f90:85-155), on the device.
//
// For each state psi: W_j = J_b T_j(H~) psi (j < N, b = x or y), P_i =
// T_i(H~) J_x psi (i < N) and mu(i, j) = g_i g_j <P_i | W_j> with the Jackson
// weights g (kpm.f90:8-29).  The Chebyshev vectors come from the term
// kernel (T_j = 2 H~ T_{j-1} - T_{j-2}, in place over T_{j-2}); the
// currents from the slice-ELL SpMV; the N x N Gram matrix P^H W is formed
// per block of up to 256 rows of mu on the FP64 matrix cores (k_gram: a
// complex product as four real 16x16x4 MFMAs, the long n dimension split
// over workgroups and summed in a fixed order by k_gram_sum).  Complex128
// only: the moments are sums of N^2 products that complex64 would leave at
// ~1e-7 of their size.
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

// Gram matrix of the Hall moments on the FP64 matrix cores:
// part[z][m][j] = sum_{k in split z} conj(P[m][k]) W[j][k] for m < rows,
// j < N, with P and W column-major (column m of P at P + m * np).  A
// workgroup (4 waves) owns a 64 x 64 tile of the output and one split of
// k; it stages 16 rows of k of both operands in LDS (real and imaginary
// parts apart, each column's 16 values together) and each wave accumulates
// a 32 x 32 sub-tile as 2 x 2 blocks of v_mfma_f64_16x16x4f64:
// re += Pr Wr + Pi Wi, im += Pr Wi - Pi Wr.
// Operand layout of the instruction: lane i + 16 k holds A[i][k] and
// B[k][i]; the result D[i][j] sits in lane j + 16 (i % 4), element i / 4
// (checked on the device with known operands).
constexpr int kGramTile = 64;   // output rows / columns per workgroup
constexpr int kGramK = 16;      // rows of k per LDS stage
constexpr int kGramLd = kGramK + 1;   // LDS stride in doubles from one operand
                                       // column to the next (its 16 values of k,
                                       // one pad): 34 banks, so the 16 lanes of
                                       // an MFMA operand read land on distinct
                                       // banks
typedef double gram_d4 __attribute__((ext_vector_type(4)));

__global__ void __launch_bounds__(256)
k_gram(const double2* __restrict__ P, const double2* __restrict__ W, int64_t np, int rows,
       int N, int64_t kchunk, double2* __restrict__ part) {
    __shared__ double lds[4][kGramTile][kGramLd];   // Pr, Pi, Wr, Wi: [column][k]
    const int m0 = blockIdx.y * kGramTile, j0 = blockIdx.x * kGramTile;
    const int64_t kb = (int64_t)blockIdx.z * kchunk;
    const int64_t ke = kb + kchunk < np ? kb + kchunk : np;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int wm = (w >> 1) * 32, wn = (w & 1) * 32;
    const int li = lane & 15, lk = lane >> 4;
    gram_d4 cre[2][2], cim[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            cre[a][b] = gram_d4{0.0, 0.0, 0.0, 0.0};
            cim[a][b] = gram_d4{0.0, 0.0, 0.0, 0.0};
        }
    // Stage k0's operands are loaded into registers while the MFMAs of the
    // previous stage run, then stored to LDS.
    const int sk = t & 15;
    double2 pq[4], xq[4];
    auto fetch = [&](int64_t k0) {
        const int64_t k = k0 + sk;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int sm = (t >> 4) + 16 * q;
            pq[q] = make_double2(0.0, 0.0);
            xq[q] = make_double2(0.0, 0.0);
            if (k < ke && m0 + sm < rows) pq[q] = P[(int64_t)(m0 + sm) * np + k];
            if (k < ke && j0 + sm < N) xq[q] = W[(int64_t)(j0 + sm) * np + k];
        }
    };
    if (kb < ke) fetch(kb);
    for (int64_t k0 = kb; k0 < ke; k0 += kGramK) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int sm = (t >> 4) + 16 * q;
            lds[0][sm][sk] = pq[q].x;
            lds[1][sm][sk] = pq[q].y;
            lds[2][sm][sk] = xq[q].x;
            lds[3][sm][sk] = xq[q].y;
        }
        __syncthreads();
        if (k0 + kGramK < ke) fetch(k0 + kGramK);
#pragma unroll
        for (int kk = 0; kk < kGramK; kk += 4) {
            double ar[2], ai[2], br[2], bi[2];
#pragma unroll
            for (int a = 0; a < 2; ++a) {
                ar[a] = lds[0][wm + 16 * a + li][kk + lk];
                ai[a] = lds[1][wm + 16 * a + li][kk + lk];
                br[a] = lds[2][wn + 16 * a + li][kk + lk];
                bi[a] = lds[3][wn + 16 * a + li][kk + lk];
            }
            // eight independent accumulators between two products into
            // the same one; each still gets its products in the order
            // re: Pr Wr, Pi Wi and im: Pr Wi, -Pi Wr
#pragma unroll
            for (int a = 0; a < 2; ++a)
#pragma unroll
                for (int b = 0; b < 2; ++b) {
                    cre[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[a], br[b], cre[a][b], 0,
                                                                     0, 0);
                    cim[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[a], bi[b], cim[a][b], 0,
                                                                     0, 0);
                }
#pragma unroll
            for (int a = 0; a < 2; ++a)
#pragma unroll
                for (int b = 0; b < 2; ++b) {
                    cre[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(ai[a], bi[b], cre[a][b], 0,
                                                                     0, 0);
                    cim[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(-ai[a], br[b], cim[a][b],
                                                                     0, 0, 0);
                }
        }
        __syncthreads();
    }
    double2* out = part + (int64_t)blockIdx.z * rows * N;
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int m = m0 + wm + 16 * a + lk + 4 * v;
                const int j = j0 + wn + 16 * b + li;
                if (m < rows && j < N)
                    out[(int64_t)m * N + j] = make_double2(cre[a][b][v], cim[a][b][v]);
            }
}

// mu[m + j * ld] = sum_z part[z][m][j] in the order z = 0, 1, ...
__global__ void k_gram_sum(const double2* __restrict__ part, int nsplit, int rows, int N,
                           double2* __restrict__ mu, int ld) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= (int64_t)rows * N) return;
    const int m = (int)(q / N), j = (int)(q % N);
    double re = 0.0, im = 0.0;
    for (int z = 0; z < nsplit; ++z) {
        const double2 v = part[(int64_t)z * rows * N + q];
        re += v.x;
        im += v.y;
    }
    mu[(int64_t)j * ld + m] = make_double2(re, im);
}

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
    // k split of the Gram blocks: enough workgroups for three per CU (the
    // registers of k_gram allow three 4-wave workgroups per CU)
    const int tiles = ((B + kGramTile - 1) / kGramTile) * ((N + kGramTile - 1) / kGramTile);
    int64_t nsplit = std::max<int64_t>(1, (3 * (int64_t)h->num_cus + tiles - 1) / tiles);
    nsplit = std::min<int64_t>(nsplit, std::max<int64_t>(1, np / (4 * kGramK)));
    const int64_t kchunk = ((np + nsplit - 1) / nsplit + kGramK - 1) / kGramK * kGramK;
    nsplit = (np + kchunk - 1) / kchunk;
    DevBuf gpart((size_t)nsplit * B * N * sizeof(double2));
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
                const dim3 grid((unsigned)((N + kGramTile - 1) / kGramTile),
                                (unsigned)((rows + kGramTile - 1) / kGramTile), (unsigned)nsplit);
                timed_launch(h, kTimeGram, [&] {
                    hipLaunchKernelGGL(k_gram, grid, dim3(256), 0, st,
                                       reinterpret_cast<const double2*>(Pb.p),
                                       reinterpret_cast<const double2*>(W.p), np, rows, N,
                                       kchunk, gpart.as<double2>());
                    TBG_HIP(hipGetLastError());
                });
                const int64_t nq = (int64_t)rows * N;
                hipLaunchKernelGGL(k_gram_sum, dim3((unsigned)((nq + 255) / 256)), dim3(256), 0,
                                   st, gpart.as<double2>(), (int)nsplit, rows, N,
                                   reinterpret_cast<double2*>(mus) + i0, N);
                TBG_HIP(hipGetLastError());
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
