// Small MLP surrogate (dense layers, GELU) trained and evaluated on the GPU:
// BF16 operands on the matrix cores (v_mfma_f32_32x32x16_bf16), FP32
// accumulation, FP32 master weights, Adam.  No reference counterpart: it
// maps disorder/strain parameters to DOS spectra produced by the TBPM path.
//
// Layout: parameters packed per layer as W_l [in][out] row-major then
// b_l [out] (FP32).  Activations are row-major [batch][features].
// Deterministic: no atomics; every reduction has a fixed order.
#include <cmath>
#include <memory>
#include <vector>

#include "tbg_internal.h"

namespace tbg {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

// float -> bf16, round to nearest even (NaN kept quiet)
__device__ inline uint16_t f2bf(float f) {
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ inline float bf2f(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }

__device__ inline float gelu(float z) { return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f)); }
__device__ inline float gelu_d(float z) {
    const float cdf = 0.5f * (1.0f + erff(z * 0.70710678118654752f));
    const float pdf = 0.39894228040143268f * expf(-0.5f * z * z);
    return cdf + z * pdf;
}

enum { EPI_STORE = 0, EPI_BIAS = 1, EPI_BIAS_GELU = 2, EPI_GELU_BWD = 3 };

struct GemmArgs {
    const uint16_t* A; int64_t sam, sak;   // A(m, k)
    const uint16_t* B; int64_t sbk, sbn;   // B(k, n)
    int M, N, K;
    float* C; int64_t scm;                 // FP32 result (pre-activation / gradient)
    uint16_t* Cb; int64_t scbm;            // BF16 result (activation / gradient)
    const float* bias;                     // [N]
    const float* zaux; int64_t szm;        // pre-activation for GELU'
    int epi;
};

// One wavefront computes a 32x32 tile of C = A B with gfx950's 32x32x16
// BF16 MFMAs; a block of 4 waves covers 64x64.  Operand lane layout of the
// instruction: A: lane l holds A[l%32][8*(l/32) + j], B: B[8*(l/32) + j][l%32],
// j<8; result register r of lane l is C[8*(r/4) + 4*(l/32) + r%4][l%32].
__global__ void __launch_bounds__(256) k_gemm_bf16(GemmArgs g) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int tm = blockIdx.y * 64 + (wv >> 1) * 32;
    const int tn = blockIdx.x * 64 + (wv & 1) * 32;
    const int li = lane & 31, lk = (lane >> 5) * 8;
    f32x16 acc;
#pragma unroll
    for (int r = 0; r < 16; ++r) acc[r] = 0.0f;
    const int m = tm + li, n = tn + li;
    for (int k0 = 0; k0 < g.K; k0 += 16) {
        bf16x8 a, b;
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const int k = k0 + lk + j;
            const uint16_t ua = (m < g.M && k < g.K) ? g.A[m * g.sam + k * g.sak] : (uint16_t)0;
            const uint16_t ub = (k < g.K && n < g.N) ? g.B[k * g.sbk + n * g.sbn] : (uint16_t)0;
            a[j] = __builtin_bit_cast(__bf16, ua);
            b[j] = __builtin_bit_cast(__bf16, ub);
        }
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
    }
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const int row = tm + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
        const int col = n;
        if (row >= g.M || col >= g.N) continue;
        float v = acc[r];
        if (g.epi == EPI_BIAS || g.epi == EPI_BIAS_GELU) v += g.bias[col];
        if (g.epi == EPI_GELU_BWD) v *= gelu_d(g.zaux[row * g.szm + col]);
        if (g.C) g.C[row * g.scm + col] = v;
        if (g.Cb) g.Cb[row * g.scbm + col] = f2bf(g.epi == EPI_BIAS_GELU ? gelu(v) : v);
    }
}

__global__ void k_to_bf16(const float* x, int64_t n, uint16_t* y) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = f2bf(x[i]);
}

// Weights of every layer to BF16 (biases stay FP32).
__global__ void k_weights_bf16(const float* P, const int64_t* offW, const int64_t* cntW, int L,
                               const int64_t* offWb, uint16_t* Wb) {
    const int l = blockIdx.y;
    if (l >= L) return;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < cntW[l];
         i += (int64_t)gridDim.x * blockDim.x)
        Wb[offWb[l] + i] = f2bf(P[offW[l] + i]);
}

// MSE loss: dY = 2 (y - t) / (B*N); per-block partial sums of (y - t)^2.
__global__ void __launch_bounds__(256) k_mse(const float* y, const float* t, int64_t n, float* dy,
                                             double* part) {
    __shared__ double sh[256];
    double s = 0.0;
    const float sc = 2.0f / (float)n;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
        const float d = y[i] - t[i];
        dy[i] = sc * d;
        s += (double)d * d;
    }
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if (threadIdx.x < k) sh[threadIdx.x] += sh[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

// Column sums of dZ [B][N] -> db [N] (fixed order over rows).
__global__ void k_colsum(const float* dz, int B, int N, float* db) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= N) return;
    float s = 0.0f;
    for (int b = 0; b < B; ++b) s += dz[(int64_t)b * N + c];
    db[c] = s;
}

__global__ void k_adam(float* p, const float* g, float* m, float* v, int64_t n, float lr, float b1,
                       float b2, float eps, float c1, float c2) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

}  // namespace tbg

using namespace tbg;

struct tbg_mlp {
    tbg_handle* h = nullptr;
    int L = 0;
    std::vector<int> w;
    std::vector<int64_t> offW, offB, offWb, cntW;
    int64_t nparam = 0, nwb = 0;
    DevBuf P, G, Mo, Ve, Wb, offW_d, cntW_d, offWb_d;
    int64_t t = 0;
    int cap = 0;
    DevBuf xb, yout, dy, part;
    std::vector<DevBuf> Z, Ab, dZ, dZb;

    void ensure(int B) {
        if (B <= cap) return;
        cap = B;
        xb.alloc((size_t)B * w[0] * 2);
        yout.alloc((size_t)B * w[L] * 4);
        dy.alloc((size_t)B * w[L] * 4);
        part.alloc(1024 * sizeof(double));
        Z.clear(); Ab.clear(); dZ.clear(); dZb.clear();
        for (int l = 0; l < L; ++l) {
            Z.emplace_back((size_t)B * w[l + 1] * 4);
            Ab.emplace_back((size_t)B * w[l + 1] * 2);
            dZ.emplace_back((size_t)B * w[l + 1] * 4);
            dZb.emplace_back((size_t)B * w[l + 1] * 2);
        }
    }

    void gemm(const GemmArgs& g) {
        dim3 grid((g.N + 63) / 64, (g.M + 63) / 64);
        hipLaunchKernelGGL(k_gemm_bf16, grid, dim3(256), 0, h->stream, g);
        TBG_HIP(hipGetLastError());
    }

    void refresh_bf16() {
        hipLaunchKernelGGL(k_weights_bf16, dim3(64, L), dim3(256), 0, h->stream, P.as<float>(),
                           offW_d.as<int64_t>(), cntW_d.as<int64_t>(), L, offWb_d.as<int64_t>(),
                           Wb.as<uint16_t>());
        TBG_HIP(hipGetLastError());
    }

    // Forward pass; hidden layers keep pre-activations (Z) and BF16
    // activations (Ab); the output layer writes yout.
    void forward(const float* x, int B) {
        ensure(B);
        const int64_t nx = (int64_t)B * w[0];
        hipLaunchKernelGGL(k_to_bf16, dim3((unsigned)((nx + 255) / 256)), dim3(256), 0, h->stream,
                           x, nx, xb.as<uint16_t>());
        TBG_HIP(hipGetLastError());
        for (int l = 0; l < L; ++l) {
            GemmArgs g{};
            g.A = l == 0 ? xb.as<uint16_t>() : Ab[l - 1].as<uint16_t>();
            g.sam = w[l]; g.sak = 1;
            g.B = Wb.as<uint16_t>() + offWb[l];
            g.sbk = w[l + 1]; g.sbn = 1;
            g.M = B; g.N = w[l + 1]; g.K = w[l];
            g.bias = P.as<float>() + offB[l];
            if (l < L - 1) {
                g.epi = EPI_BIAS_GELU;
                g.C = Z[l].as<float>(); g.scm = w[l + 1];
                g.Cb = Ab[l].as<uint16_t>(); g.scbm = w[l + 1];
            } else {
                g.epi = EPI_BIAS;
                g.C = yout.as<float>(); g.scm = w[l + 1];
            }
            gemm(g);
        }
    }

    // Backward from dy (FP32 [B][out]) into G.
    void backward(int B) {
        const int64_t ny = (int64_t)B * w[L];
        // output layer is linear: dZ_{L-1} = dY
        TBG_HIP(hipMemcpyAsync(dZ[L - 1].p, dy.p, ny * 4, hipMemcpyDeviceToDevice, h->stream));
        hipLaunchKernelGGL(k_to_bf16, dim3((unsigned)((ny + 255) / 256)), dim3(256), 0, h->stream,
                           dy.as<float>(), ny, dZb[L - 1].as<uint16_t>());
        TBG_HIP(hipGetLastError());
        for (int l = L - 1; l >= 0; --l) {
            const uint16_t* Ain = l == 0 ? xb.as<uint16_t>() : Ab[l - 1].as<uint16_t>();
            // dW_l = A_in^T dZ_l : (w[l] x B) (B x w[l+1])
            GemmArgs g{};
            g.A = Ain; g.sam = 1; g.sak = w[l];
            g.B = dZb[l].as<uint16_t>(); g.sbk = w[l + 1]; g.sbn = 1;
            g.M = w[l]; g.N = w[l + 1]; g.K = B;
            g.C = G.as<float>() + offW[l]; g.scm = w[l + 1];
            g.epi = EPI_STORE;
            gemm(g);
            hipLaunchKernelGGL(k_colsum, dim3((w[l + 1] + 255) / 256), dim3(256), 0, h->stream,
                               dZ[l].as<float>(), B, w[l + 1], G.as<float>() + offB[l]);
            TBG_HIP(hipGetLastError());
            if (l > 0) {
                // dZ_{l-1} = (dZ_l W_l^T) * gelu'(Z_{l-1})
                GemmArgs d{};
                d.A = dZb[l].as<uint16_t>(); d.sam = w[l + 1]; d.sak = 1;
                d.B = Wb.as<uint16_t>() + offWb[l]; d.sbk = 1; d.sbn = w[l + 1];
                d.M = B; d.N = w[l]; d.K = w[l + 1];
                d.C = dZ[l - 1].as<float>(); d.scm = w[l];
                d.Cb = dZb[l - 1].as<uint16_t>(); d.scbm = w[l];
                d.zaux = Z[l - 1].as<float>(); d.szm = w[l];
                d.epi = EPI_GELU_BWD;
                gemm(d);
            }
        }
    }

    double loss(const float* target, int B) {
        const int64_t ny = (int64_t)B * w[L];
        const int nb = (int)std::min<int64_t>(1024, (ny + 255) / 256);
        hipLaunchKernelGGL(k_mse, dim3(nb), dim3(256), 0, h->stream, yout.as<float>(), target, ny,
                           dy.as<float>(), part.as<double>());
        TBG_HIP(hipGetLastError());
        std::vector<double> hp(nb);
        TBG_HIP(hipMemcpyAsync(hp.data(), part.p, nb * sizeof(double), hipMemcpyDeviceToHost,
                               h->stream));
        TBG_HIP(hipStreamSynchronize(h->stream));
        double s = 0.0;
        for (double v : hp) s += v;
        return s / (double)ny;
    }
};

namespace tbg {
template <typename F>
static int mguard(F&& f) {
    try {
        f();
        return TBG_OK;
    } catch (const Error& e) {
        set_error(e.msg);
        return e.code;
    } catch (const std::exception& e) {
        set_error(e.what());
        return TBG_EINVAL;
    }
}
}  // namespace tbg

extern "C" {

int tbg_mlp_create(tbg_handle* h, int n_layers, const int* widths, tbg_mlp** out) {
    return mguard([&] {
        TBG_REQUIRE(h && widths && out, "null argument");
        TBG_REQUIRE(n_layers >= 1 && n_layers <= 16, "1..16 layers");
        TBG_HIP(hipSetDevice(h->device));
        std::unique_ptr<tbg_mlp> m(new tbg_mlp());
        m->h = h;
        m->L = n_layers;
        m->w.assign(widths, widths + n_layers + 1);
        for (int v : m->w) TBG_REQUIRE(v >= 1 && v <= 65536, "layer width out of range");
        int64_t off = 0, offb = 0;
        for (int l = 0; l < n_layers; ++l) {
            m->offW.push_back(off);
            m->cntW.push_back((int64_t)m->w[l] * m->w[l + 1]);
            off += (int64_t)m->w[l] * m->w[l + 1];
            m->offB.push_back(off);
            off += m->w[l + 1];
            m->offWb.push_back(offb);
            offb += (int64_t)m->w[l] * m->w[l + 1];
        }
        m->nparam = off;
        m->nwb = offb;
        const size_t pb = (size_t)off * 4;
        m->P.alloc(pb); m->G.alloc(pb); m->Mo.alloc(pb); m->Ve.alloc(pb);
        m->Wb.alloc((size_t)offb * 2);
        for (DevBuf* b : {&m->P, &m->G, &m->Mo, &m->Ve})
            TBG_HIP(hipMemsetAsync(b->p, 0, pb, h->stream));
        m->offW_d.alloc(n_layers * 8); m->cntW_d.alloc(n_layers * 8); m->offWb_d.alloc(n_layers * 8);
        TBG_HIP(hipMemcpyAsync(m->offW_d.p, m->offW.data(), n_layers * 8, hipMemcpyHostToDevice, h->stream));
        TBG_HIP(hipMemcpyAsync(m->cntW_d.p, m->cntW.data(), n_layers * 8, hipMemcpyHostToDevice, h->stream));
        TBG_HIP(hipMemcpyAsync(m->offWb_d.p, m->offWb.data(), n_layers * 8, hipMemcpyHostToDevice, h->stream));
        m->refresh_bf16();
        TBG_HIP(hipStreamSynchronize(h->stream));
        *out = m.release();
    });
}

int tbg_mlp_destroy(tbg_mlp* m) {
    return mguard([&] {
        if (!m) return;
        (void)hipSetDevice(m->h->device);
        delete m;
    });
}

int tbg_mlp_num_params(const tbg_mlp* m, int64_t* n) {
    return mguard([&] {
        TBG_REQUIRE(m && n, "null argument");
        *n = m->nparam;
    });
}

// Host FP32 parameters in, packed per layer (W [in][out], then b [out]);
// resets the Adam state.
int tbg_mlp_set_params(tbg_mlp* m, const float* params, int64_t n) {
    return mguard([&] {
        TBG_REQUIRE(m && params, "null argument");
        TBG_REQUIRE(n == m->nparam, "parameter count mismatch");
        TBG_HIP(hipSetDevice(m->h->device));
        TBG_HIP(hipMemcpyAsync(m->P.p, params, n * 4, hipMemcpyHostToDevice, m->h->stream));
        TBG_HIP(hipMemsetAsync(m->Mo.p, 0, n * 4, m->h->stream));
        TBG_HIP(hipMemsetAsync(m->Ve.p, 0, n * 4, m->h->stream));
        m->t = 0;
        m->refresh_bf16();
        TBG_HIP(hipStreamSynchronize(m->h->stream));
    });
}

int tbg_mlp_get_params(const tbg_mlp* m, float* params, int64_t n) {
    return mguard([&] {
        TBG_REQUIRE(m && params, "null argument");
        TBG_REQUIRE(n == m->nparam, "parameter count mismatch");
        TBG_HIP(hipMemcpyAsync(params, m->P.p, n * 4, hipMemcpyDeviceToHost, m->h->stream));
        TBG_HIP(hipStreamSynchronize(m->h->stream));
    });
}

// y [B][out] = MLP(x [B][in]); device pointers.
int tbg_mlp_forward(tbg_mlp* m, const float* x, int B, float* y) {
    return mguard([&] {
        TBG_REQUIRE(m && x && y && B >= 1, "bad argument");
        TBG_HIP(hipSetDevice(m->h->device));
        m->forward(x, B);
        TBG_HIP(hipMemcpyAsync(y, m->yout.p, (size_t)B * m->w[m->L] * 4, hipMemcpyDeviceToDevice,
                               m->h->stream));
        TBG_HIP(hipStreamSynchronize(m->h->stream));
    });
}

// MSE loss and its gradient with respect to every parameter (host FP32,
// packed like the parameters); no update.  For gradient checks.
int tbg_mlp_gradients(tbg_mlp* m, const float* x, const float* target, int B, float* grads,
                      double* loss) {
    return mguard([&] {
        TBG_REQUIRE(m && x && target && B >= 1, "bad argument");
        TBG_HIP(hipSetDevice(m->h->device));
        m->forward(x, B);
        const double l = m->loss(target, B);
        m->backward(B);
        if (grads)
            TBG_HIP(hipMemcpyAsync(grads, m->G.p, m->nparam * 4, hipMemcpyDeviceToHost,
                                   m->h->stream));
        TBG_HIP(hipStreamSynchronize(m->h->stream));
        if (loss) *loss = l;
    });
}

// One Adam step on the MSE loss of (x, target); returns the loss before
// the step.  steps > 1 repeats on the same batch.
int tbg_mlp_train(tbg_mlp* m, const float* x, const float* target, int B, int steps, float lr,
                  float beta1, float beta2, float eps, double* loss) {
    return mguard([&] {
        TBG_REQUIRE(m && x && target && B >= 1 && steps >= 1, "bad argument");
        TBG_HIP(hipSetDevice(m->h->device));
        double l = 0.0;
        for (int s = 0; s < steps; ++s) {
            m->forward(x, B);
            l = m->loss(target, B);
            m->backward(B);
            m->t += 1;
            const float c1 = 1.0f - std::pow(beta1, (float)m->t);
            const float c2 = 1.0f - std::pow(beta2, (float)m->t);
            hipLaunchKernelGGL(k_adam, dim3((unsigned)((m->nparam + 255) / 256)), dim3(256), 0,
                               m->h->stream, m->P.as<float>(), m->G.as<float>(), m->Mo.as<float>(),
                               m->Ve.as<float>(), m->nparam, lr, beta1, beta2, eps, c1, c2);
            TBG_HIP(hipGetLastError());
            m->refresh_bf16();
        }
        TBG_HIP(hipStreamSynchronize(m->h->stream));
        if (loss) *loss = l;
    });
}

// Plain BF16 GEMM through the same MFMA kernel: C [M][N] (FP32) = A [M][K]
// B [K][N], row-major BF16 inputs (device).  For tests of the matrix path.
int tbg_gemm_bf16(tbg_handle* h, int M, int N, int K, const uint16_t* A, const uint16_t* B,
                  float* C) {
    return mguard([&] {
        TBG_REQUIRE(h && A && B && C && M >= 1 && N >= 1 && K >= 1, "bad argument");
        TBG_HIP(hipSetDevice(h->device));
        GemmArgs g{};
        g.A = A; g.sam = K; g.sak = 1;
        g.B = B; g.sbk = N; g.sbn = 1;
        g.M = M; g.N = N; g.K = K;
        g.C = C; g.scm = N;
        g.epi = EPI_STORE;
        dim3 grid((N + 63) / 64, (M + 63) / 64);
        hipLaunchKernelGGL(k_gemm_bf16, grid, dim3(256), 0, h->stream, g);
        TBG_HIP(hipGetLastError());
        TBG_HIP(hipStreamSynchronize(h->stream));
    });
}

}  // extern "C"
