// Propagation engine: the fused Chebyshev-term kernel and its drivers.
#pragma once

#include "tbg_internal.h"

namespace tbg {

// Flags of the fused term kernel (compile-time).
enum TermFlags : int {
    F_PREV = 1,      // recurrence uses beta * v_{m-2} (row-aligned read)
    F_STORE = 2,     // store v_m (off for the last term of a step)
    F_ACC = 4,       // accumulate into the step output
    F_ACCRMW = 8,    // ... starting from its current value (else from 0)
    F_ACCPREV = 16,  // ... include cp * v_{m-2}
    F_MID = 32,      // ... include cm * v_{m-1} (row-aligned read of pin)
    F_DOT = 64,      // epilogue: sum conj(psi0) * out per vector (FP64)
    F_NORM = 128,    // epilogue: sum |out|^2 per vector (FP64)
};

// One Chebyshev term over a (possibly block-diagonal, batched) sliced-ELL
// matrix: v_m = alpha * H v_{m-1} + beta * v_{m-2}, plus accumulation.
struct TermArgs {
    const int* cols;
    const void* vals;
    const int64_t* soff;     // slice element offsets
    const int* sw;           // slice widths
    const int64_t* chunk;    // [nchunks+1] slice ranges per block
    int64_t nslices;
    const void* pin;         // v_{m-1}, gathered
    const void* pprev;       // v_{m-2}, row aligned (may alias pout)
    void* pout;              // v_m
    void* acc;               // step output
    const void* psi0;        // for F_DOT
    double ar[2], ai[2];     // alpha per coefficient class
    double beta;
    double cp[2], cm[2], cc[2];
    uint64_t bwd_mask;       // vectors propagated with conj(alpha)
    uint64_t cls_mask;       // vectors using coefficient class 1
    int K, lgG;
    double* part;            // [nchunks][nv] partial sums
    int nv;                  // values per partial row (3K for dot+norm)
    int* counter;            // incremented once by block 0 when non-null
    int stage;               // stage the slice's own rows of pin in LDS
    int real_vals;           // vals holds R (all imaginary parts zero)
};

// Vectors one lane of the term kernel carries for Kp slots; the rest of a
// row is shared by G = Kp/V lanes and a wave covers its slice in G passes.
// Fewer passes put more gathers in flight per pass, more vectors per lane
// cost registers.  Chosen from sweeps on 8.4M-orbital graphene (best launch
// shape each): complex64 K=8 runs 2% faster with V=4 than V=2, K=4 and K=32
// 1-2% faster with V=2, V=8 loses 10% (occupancy 2-4 waves per SIMD);
// complex128 uses up to 4.  Slot counts 3 * 2^j (slots_for) use V = 3.
template <typename R>
inline int vectors_per_lane(int Kp) {
    if (Kp % 3 == 0) return 3;
    if (sizeof(R) == 8) return Kp < 4 ? Kp : 4;
    if (Kp == 8) return 4;
    return Kp < 2 ? Kp : 2;
}

// Blocks are dispatched round-robin over the 8 XCDs (block b -> XCD b%8).
// Give each XCD one contiguous run of chunks so the neighbour rows a chunk
// gathers were touched by the same XCD's L2.
__device__ __forceinline__ int remap_chunk(int b, int nb) {
    const int x = b & 7, i = b >> 3;
    const int q = nb >> 3, r = nb & 7;
    return x * q + (x < r ? x : r) + i;
}

// Plan of the persistent launch: chunk table + sizes.
struct Plan {
    int64_t nslices = 0;
    int nchunks = 0;
    DevBuf chunk;            // int64 [nchunks+1]
    std::vector<int64_t> chunk_host;
    std::vector<int> inst_chunk;   // [ninst+1] chunk ranges per instance
    DevBuf inst_chunk_dev;
};

// Build a plan over nslices, splitting at the slice boundaries in
// `inst_slices` (size ninst+1) so no chunk spans two instances.
void make_plan(tbg_handle* h, Plan& plan, const std::vector<int64_t>& inst_slices,
               int target_chunks);

// Kernel launchers (defined in tbg_prop.hip) -------------------------------
template <typename R> void launch_term(int flags, const TermArgs& a, int nchunks,
                                       hipStream_t s);
// Sum partial rows per instance: out[i][(*counter + toff) * nv + q].
void launch_reduce(const double* part, int nv, const int* inst_chunk, int ninst,
                   double* out, int64_t inst_stride, const int* counter, int toff,
                   hipStream_t s);
// pack complex128 (K, n) -> interleaved [n_pad][Kp] complex<R> at slot offset;
// with perm (layout row -> caller row) layout row r takes caller row perm[r].
template <typename R> void launch_pack(const double* src, int K, int64_t n,
                                       void* dst, int Kp, int slot0, const int* perm,
                                       hipStream_t s);
template <typename R> void launch_unpack(const void* src, int Kp, int slot0, int64_t n,
                                         double* dst, int K, const int* perm, hipStream_t s);
// Per-vector dot/norm of packed vectors a (conj) and b -> partials (3 per vector).
template <typename R> void launch_dotnorm(const Plan& p, const void* a, const void* b,
                                          int Kp, int K, double* part, hipStream_t s);
// y[slot] = M x[slot] for nk vectors (M = H values or a current operator).
template <typename R> void launch_spmv(const tbg_ham* H, const void* mvals, const void* x,
                                       int Kx, int sx, void* y, int Ky, int sy, int nk,
                                       hipStream_t s);
void launch_fill_zero(void* p, size_t bytes, hipStream_t s);

}  // namespace tbg
