// Device helpers shared by the propagation kernels: per-lane vectors of V
// complex states, their loads and stores, and the arithmetic every sparse
// product and recurrence step uses (one definition, so the one-term and the
// three-term kernels produce bit-identical states).
#pragma once

#include <type_traits>

#include "tbg_internal.h"

namespace tbg {

template <typename R, int V> struct Vec {
    R re[V];
    R im[V];
};

template <typename R, int V>
__device__ __forceinline__ Vec<R, V> vzero() {
    Vec<R, V> o;
#pragma unroll
    for (int v = 0; v < V; ++v) { o.re[v] = R(0); o.im[v] = R(0); }
    return o;
}

// idx is in complex elements; V consecutive complex values from idx (idx
// even when V is even: complex64 pairs are read as one 16-byte load).
template <typename R, int V>
__device__ __forceinline__ Vec<R, V> vld(const void* base, int64_t idx) {
    Vec<R, V> o;
    if constexpr (std::is_same<R, float>::value && V % 2 == 0) {
#pragma unroll
        for (int q = 0; q < V / 2; ++q) {
            const float4 t = reinterpret_cast<const float4*>(base)[(idx >> 1) + q];
            o.re[2 * q] = t.x; o.im[2 * q] = t.y; o.re[2 * q + 1] = t.z; o.im[2 * q + 1] = t.w;
        }
    } else if constexpr (std::is_same<R, float>::value) {
#pragma unroll
        for (int v = 0; v < V; ++v) {
            const float2 t = reinterpret_cast<const float2*>(base)[idx + v];
            o.re[v] = t.x; o.im[v] = t.y;
        }
    } else {
#pragma unroll
        for (int v = 0; v < V; ++v) {
            const double2 t = reinterpret_cast<const double2*>(base)[idx + v];
            o.re[v] = t.x; o.im[v] = t.y;
        }
    }
    return o;
}

template <typename R, int V>
__device__ __forceinline__ void vst(void* base, int64_t idx, const Vec<R, V>& o) {
    if constexpr (std::is_same<R, float>::value && V % 2 == 0) {
#pragma unroll
        for (int q = 0; q < V / 2; ++q)
            reinterpret_cast<float4*>(base)[(idx >> 1) + q] =
                make_float4(o.re[2 * q], o.im[2 * q], o.re[2 * q + 1], o.im[2 * q + 1]);
    } else if constexpr (std::is_same<R, float>::value) {
#pragma unroll
        for (int v = 0; v < V; ++v)
            reinterpret_cast<float2*>(base)[idx + v] = make_float2(o.re[v], o.im[v]);
    } else {
#pragma unroll
        for (int v = 0; v < V; ++v)
            reinterpret_cast<double2*>(base)[idx + v] = make_double2(o.re[v], o.im[v]);
    }
}

// Streams touched once per term (v_{m-2}, the step output, psi0, v_m) can be
// loaded and stored with the non-temporal hint, so the L2 keeps the rows of
// v_{m-1} that neighbouring slices gather.  Measured on 8.4M-orbital
// graphene this gains 8-9% with one vector per lane (complex128 or
// complex64, K = 1) and loses 10-35% with two or more (K >= 4), so the term
// kernel uses it for V = 1 only (stream_ld / stream_st).
typedef float tbg_f4 __attribute__((ext_vector_type(4)));
typedef float tbg_f2 __attribute__((ext_vector_type(2)));
typedef double tbg_d2 __attribute__((ext_vector_type(2)));

template <typename R, int V>
__device__ __forceinline__ Vec<R, V> vld_nt(const void* base, int64_t idx) {
    Vec<R, V> o;
    if constexpr (std::is_same<R, float>::value && V % 2 == 0) {
#pragma unroll
        for (int q = 0; q < V / 2; ++q) {
            const tbg_f4 t = __builtin_nontemporal_load(
                reinterpret_cast<const tbg_f4*>(base) + (idx >> 1) + q);
            o.re[2 * q] = t.x; o.im[2 * q] = t.y; o.re[2 * q + 1] = t.z; o.im[2 * q + 1] = t.w;
        }
    } else if constexpr (std::is_same<R, float>::value) {
#pragma unroll
        for (int v = 0; v < V; ++v) {
            const tbg_f2 t =
                __builtin_nontemporal_load(reinterpret_cast<const tbg_f2*>(base) + idx + v);
            o.re[v] = t.x; o.im[v] = t.y;
        }
    } else {
#pragma unroll
        for (int v = 0; v < V; ++v) {
            const tbg_d2 t = __builtin_nontemporal_load(
                reinterpret_cast<const tbg_d2*>(base) + idx + v);
            o.re[v] = t.x; o.im[v] = t.y;
        }
    }
    return o;
}

template <typename R, int V>
__device__ __forceinline__ void vst_nt(void* base, int64_t idx, const Vec<R, V>& o) {
    if constexpr (std::is_same<R, float>::value && V % 2 == 0) {
#pragma unroll
        for (int q = 0; q < V / 2; ++q) {
            const tbg_f4 t = {o.re[2 * q], o.im[2 * q], o.re[2 * q + 1], o.im[2 * q + 1]};
            __builtin_nontemporal_store(t, reinterpret_cast<tbg_f4*>(base) + (idx >> 1) + q);
        }
    } else if constexpr (std::is_same<R, float>::value) {
#pragma unroll
        for (int v = 0; v < V; ++v) {
            const tbg_f2 t = {o.re[v], o.im[v]};
            __builtin_nontemporal_store(t, reinterpret_cast<tbg_f2*>(base) + idx + v);
        }
    } else {
#pragma unroll
        for (int v = 0; v < V; ++v) {
            const tbg_d2 t = {o.re[v], o.im[v]};
            __builtin_nontemporal_store(t, reinterpret_cast<tbg_d2*>(base) + idx + v);
        }
    }
}

template <typename R, int V>
__device__ __forceinline__ Vec<R, V> stream_ld(const void* base, int64_t idx) {
    if constexpr (V == 1) return vld_nt<R, V>(base, idx);
    else return vld<R, V>(base, idx);
}
template <typename R, int V>
__device__ __forceinline__ void stream_st(void* base, int64_t idx, const Vec<R, V>& o) {
    if constexpr (V == 1) vst_nt<R, V>(base, idx, o);
    else vst<R, V>(base, idx, o);
}

// Value of one stored entry of H: complex, or real when the Hamiltonian has
// no imaginary part.
template <typename R, bool RV> struct HVal;
template <typename R> struct HVal<R, true> {
    R x;
    __device__ __forceinline__ void load(const void* vals, int e) {
        x = reinterpret_cast<const R*>(vals)[e];
    }
};
template <typename R> struct HVal<R, false> {
    typename CplxT<R>::T c;
    __device__ __forceinline__ void load(const void* vals, int e) {
        c = reinterpret_cast<const typename CplxT<R>::T*>(vals)[e];
    }
};

// s += h x (real values skip the products with an exact zero).
template <typename R, int V, bool RV>
__device__ __forceinline__ void mac(Vec<R, V>& s, const HVal<R, RV>& h, const Vec<R, V>& x) {
#pragma unroll
    for (int v = 0; v < V; ++v) {
        if constexpr (RV) {
            s.re[v] = fma(h.x, x.re[v], s.re[v]);
            s.im[v] = fma(h.x, x.im[v], s.im[v]);
        } else {
            s.re[v] = fma(h.c.x, x.re[v], s.re[v]);
            s.re[v] = fma(-h.c.y, x.im[v], s.re[v]);
            s.im[v] = fma(h.c.x, x.im[v], s.im[v]);
            s.im[v] = fma(h.c.y, x.re[v], s.im[v]);
        }
    }
}

// One recurrence step from the product s = H v_{m-1}:
// alpha s (+ beta prev when PREV), alpha = ar + i ai per state.
template <typename R, int V, bool PREV>
__device__ __forceinline__ Vec<R, V> recur(const R* ar, const R* ai, R beta, const Vec<R, V>& s,
                                           const Vec<R, V>& prev) {
    Vec<R, V> cur;
#pragma unroll
    for (int v = 0; v < V; ++v) {
        cur.re[v] = fma(ar[v], s.re[v], -(ai[v] * s.im[v]));
        cur.im[v] = fma(ar[v], s.im[v], ai[v] * s.re[v]);
        if constexpr (PREV) {
            cur.re[v] = fma(beta, prev.re[v], cur.re[v]);
            cur.im[v] = fma(beta, prev.im[v], cur.im[v]);
        }
    }
    return cur;
}

// c ? x : y per register (no branch; x and y by value, so the select stays
// one of values and a queue of these stays in registers).
template <typename R, int V>
__device__ __forceinline__ Vec<R, V> vsel(bool c, Vec<R, V> x, Vec<R, V> y) {
    Vec<R, V> o;
#pragma unroll
    for (int v = 0; v < V; ++v) {
        o.re[v] = c ? x.re[v] : y.re[v];
        o.im[v] = c ? x.im[v] : y.im[v];
    }
    return o;
}

// o += c x per state.
template <typename R, int V>
__device__ __forceinline__ void axpy(Vec<R, V>& o, const R* c, const Vec<R, V>& x) {
#pragma unroll
    for (int v = 0; v < V; ++v) {
        o.re[v] = fma(c[v], x.re[v], o.re[v]);
        o.im[v] = fma(c[v], x.im[v], o.im[v]);
    }
}

// s += sum_{j0 <= j < j0+NJ} H[row][j] x[idx_j]: the indices and values of
// the NJ entries are loaded first, then all NJ gathers are issued, then the
// products run in j order, so NJ gathers are in flight at once (a loop over
// j would wait for each index before its gather).  idx(j) loads the index
// of entry j, ent(j) its element offset from vals (an int: vals points at
// the row's slice), gather(i) the states at index i.
template <typename R, int V, bool RV, int NJ, typename IDX, typename ENT, typename GATHER>
__device__ __forceinline__ void sum_group(Vec<R, V>& s, const void* vals, int j0, IDX idx,
                                          ENT ent, GATHER gather) {
    int col[NJ];
    HVal<R, RV> h[NJ];
#pragma unroll
    for (int u = 0; u < NJ; ++u) {
        col[u] = idx(j0 + u);
        h[u].load(vals, ent(j0 + u));
    }
    Vec<R, V> x[NJ];
#pragma unroll
    for (int u = 0; u < NJ; ++u) x[u] = gather(col[u]);
#pragma unroll
    for (int u = 0; u < NJ; ++u) mac<R, V, RV>(s, h[u], x[u]);
}

// The whole row (w entries) in groups of up to four.
template <typename R, int V, bool RV, typename IDX, typename ENT, typename GATHER>
__device__ __forceinline__ Vec<R, V> row_sum(int w, const void* vals, IDX idx, ENT ent,
                                             GATHER gather) {
    Vec<R, V> s = vzero<R, V>();
    int j0 = 0;
    for (; j0 + 4 <= w; j0 += 4) sum_group<R, V, RV, 4>(s, vals, j0, idx, ent, gather);
    switch (w - j0) {
        case 1: sum_group<R, V, RV, 1>(s, vals, j0, idx, ent, gather); break;
        case 2: sum_group<R, V, RV, 2>(s, vals, j0, idx, ent, gather); break;
        case 3: sum_group<R, V, RV, 3>(s, vals, j0, idx, ent, gather); break;
        default: break;
    }
    return s;
}

}  // namespace tbg
