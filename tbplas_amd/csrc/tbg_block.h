// Temporal blocking: two to four Chebyshev terms per pass over a spatial
// tile of rows (tbg_block.hip).
#pragma once

#include <memory>
#include <vector>

#include "tbg_prop.h"

namespace tbg {

// Rows of one block-tile: R = layout rows [r0, r0 + n[0]) (a run of the
// caller's spatial tiles), halo ring E1 = neighbours of R outside R, ring
// E(k+1) = neighbours of Ek outside R and E1..Ek, for depth rings.
// rows[offA ..] lists R, E1, E2, .. in that order; n[k] = rows of R and the
// first k rings (a prefix; n[k] = n[depth] beyond the depth).  A = all
// nA = n[depth] rows, B = the first nB = n[depth - 1].  Every row of A has w
// entries (the widest row of A; shorter rows padded with zero values on
// their own column), stored column-major so a workgroup reads them
// coalesced: colG[offG + j*nA + l] (layout column), valG[offG + j*nA + l]
// (value, the layout's element type) and, for the rows of B,
// colL[offL + j*nB + l] (position of that column in A, 16 bits: A fits the
// LDS, so it has fewer than 65536 rows).  rows[offA + l] = r0 + l for the
// owned rows l < n[0].
struct BlockMeta {
    int64_t offA, offL, offG;
    int n[5];
    int nA, nB, w, r0;
};

struct Blocking {
    int row_bytes = 0;       // bytes of one row of states it was sized for
    int depth = 2;           // halo rings (2 to 4): passes of up to depth + 1 terms
    int nblocks = 0;
    std::vector<int> seg_block;   // blocks of segment i: [seg_block[i], seg_block[i+1])
    int nA_max = 0, nB_max = 0;
    double halo = 0.0;       // (sum nA) / n: rows computed per row owned, term m
    // What a full pass of depth + 1 terms must move, summed over the tiles:
    // table bytes, distinct rows of v_{m-1} gathered, rows of v_{m-2} read
    // (sum nA), owned rows (sum n[0]).
    int64_t sched[4] = {0, 0, 0, 0};
    DevBuf meta;             // BlockMeta [nblocks]
    DevBuf rows;             // int32 [sum nA]
    DevBuf colL;             // uint16 [sum w * nB]
    DevBuf colG;             // int32 [sum w * nA]
    DevBuf valG;             // value [sum w * nA]
};

// The blocking of H for rows of row_bytes bytes and the handle's terms per
// pass (built on first use and kept on H), or null when H has no tiles,
// blocking is off on the handle, or a block-tile would not fit the LDS of
// one workgroup.
std::shared_ptr<Blocking> get_blocking(tbg_handle* h, const tbg_ham* H, int row_bytes);

// A sliced-ELL matrix on the device (one Hamiltonian, or the
// block-diagonal concatenation of a batch): nrows rows in the index space.
struct BlockSrc {
    const int* cols = nullptr;
    const void* vals = nullptr;
    const int* sw = nullptr;
    const int64_t* soff = nullptr;
    int64_t nslices = 0, padded = 0, nrows = 0;
    size_t vsz = 0;            // bytes per stored value
};

// Tables with depth (2 to 4) halo rings for tiles given per segment
// (absolute row offsets, first = the segment's start, last = its end);
// blocks never cross a segment end; null when a block-tile would not fit
// the LDS of one workgroup.
std::shared_ptr<Blocking> build_blocking(tbg_handle* h, const BlockSrc& src,
                                         const std::vector<std::vector<int64_t>>& segs,
                                         int row_bytes, int depth);

// Kinds of pass over the block-tiles (k_block_pass, T terms m .. m+T-1):
//   PM_NORMAL  acc += sum_t c_{m+t} v_{m+t}; the last two terms to q1, q2
//   PM_M1      m = 1 from v_0 (p1): v_1 = alpha_1 H v_0 (ar1, ai1), no v_{-1};
//              acc = c_0 v_0 + sum_t c_{1+t} v_{1+t}; bumps *counter as the
//              one-term kernel's m = 1 launch does
//   PM_LAST, PM_NORM, PM_DOTNORM  the step's last terms: acc += sum_t
//              c_{m+t} v_{m+t}, nothing else written, with no epilogue, the
//              FP64 norm or the FP64 dot with psi0 and the norm summed per
//              tile into part[tile][nv] (layout of the one-term kernel's
//              F_DOT / F_NORM partial rows).  The only modes of T = 2.
enum PassMode : int { PM_NORMAL = 0, PM_M1, PM_LAST, PM_NORM, PM_DOTNORM };

// Arguments of the tile-pass kernel: from v_{m-2} (p0) and v_{m-1} (p1) it
// forms v_m .. v_{m+T-1} on each block-tile (all but the last on halo
// rings, in LDS), writes the last two to q1, q2 for the owned rows (unless
// the pass ends the step), and updates the step output acc (see PassMode).
struct PassArgs {
    const BlockMeta* meta;
    const int* rows;
    const uint16_t* colL;
    const int* colG;
    const void* valG;
    const void* p0;
    const void* p1;
    void* q1;
    void* q2;
    void* acc;
    double ar[2], ai[2];
    double ar1[2], ai1[2];   // alpha of the first term (PM_M1)
    double beta;
    double c[7][2];          // c_{m-2} .. c_{m+4} per coefficient class
    uint64_t bwd_mask, cls_mask;
    int K, lgG;
    int nA_max, nB_max;      // widest A and B: B starts at row nA_max of the LDS,
                             // its spare row (phase 1) is row nB_max of B
    int* counter;            // PM_M1
    const void* psi0;        // PM_DOTNORM
    double* part;            // PM_NORM, PM_DOTNORM
    int nv;                  // values per partial row (3 K)
};

// States per lane of the tile-pass kernel for K state slots: 3 for
// multiples of 3, else 2 (1 for one slot), independent of the element type.
// More lanes per row than the one-term kernel uses (which carries up to 4)
// keep more gathers in flight per tile: complex64 K=8 +5.7%, complex128
// K=8 +11%, complex64 K=16 +41% (8.4M-orbital graphene, 3.5 nm tiles).
inline int term3_vectors_per_lane(int K) { return K % 3 == 0 ? 3 : (K < 2 ? K : 2); }

// Workgroup size of the tile-pass kernel for K slots of elem_bytes-byte
// reals (4 or 8), real or complex values, and tiles of at most nA rows.
int term3_launch_threads(int elem_bytes, bool real_vals, int K, int nA);

// One pass of T terms (2 <= T <= b.depth + 1); a.lgG = log2(K /
// term3_vectors_per_lane(K)).
template <typename R>
void launch_pass(int mode, int T, bool real_vals, const PassArgs& a, const Blocking& b,
                 hipStream_t s);

// Lets the tile-pass kernels use max_lds bytes of dynamic LDS (per device,
// once, before any launch).
void prepare_term3(int max_lds);

// The instances of real type R with T terms per pass and real (RV) or
// complex values (tbg_block_pass.h, instantiated in tbg_pass_*.hip):
// launch, and the LDS permission.
template <typename R, int T, bool RV>
void launch_pass_rt(int mode, const PassArgs& a, const Blocking& b, hipStream_t s);
template <typename R, int T, bool RV>
void allow_lds_rt(int bytes);

}  // namespace tbg
