// Supercell hop lists on the device: build_hop_pbc (tbplas/cython/
// super.pyx:860-986) and build_hop_gen (super.pyx:643-783) with the index
// helpers _id_pc2sc (19-43), _id_pc2sc_vac2 (104-150), _wrap_pbc (446-494)
// and _fast_div (496-539), combined as SuperCell.get_hop does without a hop
// modifier (tbplas/builder/super.py:703-741, 743-775).
//
// build_hop_pbc emits one term per (cell, periodic hop) at a position known
// in advance; build_hop_gen keeps the terms that stay inside open
// boundaries and off vacancies, in (orbital, hop) order: a count per
// orbital, an exclusive scan (hipCUB) and a fill keep that order.  The
// distances add the terms in the reference's order and without fused
// multiply-add, so they agree bit for bit.
#pragma clang fp contract(off)

#include <hipcub/hipcub.hpp>

#include <vector>

#include "tbg_internal.h"

namespace tbg {

namespace {

struct ScGeom {
    int d0, d1, d2;
    int p0, p1, p2;
    int norb_pc;
    double L[9];   // supercell lattice vectors, row-major (3, 3)
};

__device__ __forceinline__ int wrap_pbc(int j, int n, int pbc) {
    if (pbc == 0) return (j < 0 || j >= n) ? -1 : j;
    int r = j % n;
    return r < 0 ? r + n : r;
}

__device__ __forceinline__ int floor_div(int j, int n) {
    int q = j / n;
    if (j % n != 0 && j < 0) q -= 1;
    return q;
}

__device__ __forceinline__ int64_t pc2sc(const ScGeom& g, int a, int b, int c, int o) {
    const int64_t cell = (int64_t)a * g.d1 * g.d2 + (int64_t)b * g.d2 + c;
    return cell * g.norb_pc + o;
}

// id in the supercell without vacancies -> id with them removed, -1 for a
// vacancy (vac sorted ascending)
__device__ __forceinline__ int64_t skip_vac(int64_t id, const int64_t* vac, int64_t nv) {
    int64_t lo = 0, hi = nv;   // count of vac < id
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if (vac[mid] < id) lo = mid + 1;
        else hi = mid;
    }
    if (lo < nv && vac[lo] == id) return -1;
    return id - lo;
}

__device__ __forceinline__ void put_dr(double* dr, int64_t p, const double* pos, int64_t i,
                                       int64_t j, int na, int nb, int nc, const ScGeom& g) {
    for (int x = 0; x < 3; ++x) {
        double v = pos[j * 3 + x] - pos[i * 3 + x];
        v = v + (double)na * g.L[0 * 3 + x];
        v = v + (double)nb * g.L[1 * 3 + x];
        v = v + (double)nc * g.L[2 * 3 + x];
        dr[p * 3 + x] = v;
    }
}

// build_hop_pbc: term (cell, ih) at base + cell * nh + ih
__global__ void k_hop_pbc(const int* __restrict__ ind, const double2* __restrict__ eng, int nh,
                          ScGeom g, const double* __restrict__ pos, int64_t base,
                          int64_t* __restrict__ hi, int64_t* __restrict__ hj,
                          double2* __restrict__ hv, double* __restrict__ dr) {
    const int64_t ncell = (int64_t)g.d0 * g.d1 * g.d2;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ncell * nh) return;
    const int64_t cell = t / nh;
    const int ih = (int)(t - cell * nh);
    const int ia = (int)(cell / ((int64_t)g.d1 * g.d2));
    const int ib = (int)((cell / g.d2) % g.d1);
    const int ic = (int)(cell % g.d2);
    const int* h = ind + ih * 5;
    const int64_t oi = pc2sc(g, ia, ib, ic, 0);
    int64_t oj = oi;
    int na = 0, nb = 0, nc = 0;
    if (!(h[0] == 0 && h[1] == 0 && h[2] == 0)) {
        const int ja0 = ia + h[0], jb0 = ib + h[1], jc0 = ic + h[2];
        oj = pc2sc(g, wrap_pbc(ja0, g.d0, g.p0), wrap_pbc(jb0, g.d1, g.p1),
                   wrap_pbc(jc0, g.d2, g.p2), 0);
        na = floor_div(ja0, g.d0);
        nb = floor_div(jb0, g.d1);
        nc = floor_div(jc0, g.d2);
    }
    const int64_t i = oi + h[3], j = oj + h[4], p = base + t;
    hi[p] = i;
    hj[p] = j;
    hv[p] = eng[ih];
    put_dr(dr, p, pos, i, j, na, nb, nc, g);
}

// build_hop_gen for orbital io and hop ih: the ket id, or -1 when dropped
__device__ __forceinline__ int64_t gen_target(const int* o, const int* h, const ScGeom& g,
                                              const int64_t* vac, int64_t nv, int& na, int& nb,
                                              int& nc) {
    if (o[3] != h[3]) return -1;
    const int ja0 = o[0] + h[0], jb0 = o[1] + h[1], jc0 = o[2] + h[2];
    const int ja = wrap_pbc(ja0, g.d0, g.p0);
    const int jb = wrap_pbc(jb0, g.d1, g.p1);
    const int jc = wrap_pbc(jc0, g.d2, g.p2);
    if (ja == -1 || jb == -1 || jc == -1) return -1;
    int64_t j = pc2sc(g, ja, jb, jc, h[4]);
    if (nv > 0) j = skip_vac(j, vac, nv);
    na = floor_div(ja0, g.d0);
    nb = floor_div(jb0, g.d1);
    nc = floor_div(jc0, g.d2);
    return j;
}

__global__ void k_hop_gen_count(const int* __restrict__ orb, int64_t norb,
                                const int* __restrict__ ind, int nh, ScGeom g,
                                const int64_t* __restrict__ vac, int64_t nv,
                                int64_t* __restrict__ cnt) {
    const int64_t io = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (io >= norb) return;
    int64_t c = 0;
    int na, nb, nc;
    for (int ih = 0; ih < nh; ++ih)
        if (gen_target(orb + io * 4, ind + ih * 5, g, vac, nv, na, nb, nc) != -1) ++c;
    cnt[io] = c;
}

__global__ void k_hop_gen_fill(const int* __restrict__ orb, int64_t norb,
                               const int* __restrict__ ind, const double2* __restrict__ eng,
                               int nh, ScGeom g, const int64_t* __restrict__ vac, int64_t nv,
                               const int64_t* __restrict__ off, const double* __restrict__ pos,
                               int64_t base, int64_t* __restrict__ hi, int64_t* __restrict__ hj,
                               double2* __restrict__ hv, double* __restrict__ dr) {
    const int64_t io = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (io >= norb) return;
    int64_t p = base + off[io];
    int na, nb, nc;
    for (int ih = 0; ih < nh; ++ih) {
        const int64_t j = gen_target(orb + io * 4, ind + ih * 5, g, vac, nv, na, nb, nc);
        if (j == -1) continue;
        hi[p] = io;
        hj[p] = j;
        hv[p] = eng[ih];
        put_dr(dr, p, pos, io, j, na, nb, nc, g);
        ++p;
    }
}

unsigned blocks(int64_t n) { return (unsigned)((n + 255) / 256); }

}  // namespace

// All arrays are host arrays; the work runs on h's device.  Returns the
// number of terms written (at most capacity).
int64_t supercell_hops(tbg_handle* h, const int* pc_hop_ind, const double* pc_hop_eng, int nh,
                       const int* dim, const int* pbc, int norb_pc, const int* orb_id_pc,
                       int64_t norb_sc, const int64_t* vac, int64_t nvac, const double* sc_lat,
                       const double* sc_pos, int64_t capacity, int64_t* hop_i, int64_t* hop_j,
                       double* hop_v, double* dr) {
    const hipStream_t st = h->stream;
    ScGeom g{dim[0], dim[1], dim[2], pbc[0] != 0, pbc[1] != 0, pbc[2] != 0, norb_pc, {}};
    for (int x = 0; x < 9; ++x) g.L[x] = sc_lat[x];
    const int64_t ncell = (int64_t)dim[0] * dim[1] * dim[2];
    // split_pc_hop (super.pyx:786-857): periodic = no offset along an open
    // direction; used only without vacancies (super.py:770-773)
    std::vector<int> ip, ifr;
    std::vector<double> ep, efr;
    for (int k = 0; k < nh; ++k) {
        bool per = true;
        for (int x = 0; x < 3; ++x)
            if (pbc[x] == 0 && pc_hop_ind[k * 5 + x] != 0) per = false;
        auto& I = (nvac == 0 && per) ? ip : ifr;
        auto& E = (nvac == 0 && per) ? ep : efr;
        I.insert(I.end(), pc_hop_ind + k * 5, pc_hop_ind + k * 5 + 5);
        E.push_back(pc_hop_eng[2 * k]);
        E.push_back(pc_hop_eng[2 * k + 1]);
    }
    const int nhp = (int)ep.size() / 2, nhf = (int)efr.size() / 2;
    TBG_REQUIRE(ncell * nhp <= capacity, "output capacity too small");
    DevBuf dpos((size_t)norb_sc * 3 * sizeof(double));
    TBG_HIP(hipMemcpyAsync(dpos.p, sc_pos, dpos.bytes, hipMemcpyHostToDevice, st));
    const int64_t cap = std::max<int64_t>(capacity, 1);
    DevBuf dhi(cap * 8), dhj(cap * 8), dhv(cap * 16), ddr(cap * 24);
    int64_t total = 0;
    if (nhp > 0) {
        DevBuf dind(ip.size() * sizeof(int)), deng(ep.size() * sizeof(double));
        TBG_HIP(hipMemcpyAsync(dind.p, ip.data(), dind.bytes, hipMemcpyHostToDevice, st));
        TBG_HIP(hipMemcpyAsync(deng.p, ep.data(), deng.bytes, hipMemcpyHostToDevice, st));
        hipLaunchKernelGGL(k_hop_pbc, dim3(blocks(ncell * nhp)), dim3(256), 0, st,
                           dind.as<int>(), deng.as<double2>(), nhp, g, dpos.as<double>(),
                           (int64_t)0, dhi.as<int64_t>(), dhj.as<int64_t>(), dhv.as<double2>(),
                           ddr.as<double>());
        TBG_HIP(hipGetLastError());
        TBG_HIP(hipStreamSynchronize(st));
        total = ncell * nhp;
    }
    if (nhf > 0 && norb_sc > 0) {
        DevBuf dind(ifr.size() * sizeof(int)), deng(efr.size() * sizeof(double));
        DevBuf dorb((size_t)norb_sc * 4 * sizeof(int));
        DevBuf dvac((size_t)std::max<int64_t>(nvac, 1) * sizeof(int64_t));
        DevBuf cnt((size_t)norb_sc * sizeof(int64_t)), off((size_t)norb_sc * sizeof(int64_t));
        TBG_HIP(hipMemcpyAsync(dind.p, ifr.data(), dind.bytes, hipMemcpyHostToDevice, st));
        TBG_HIP(hipMemcpyAsync(deng.p, efr.data(), deng.bytes, hipMemcpyHostToDevice, st));
        TBG_HIP(hipMemcpyAsync(dorb.p, orb_id_pc, dorb.bytes, hipMemcpyHostToDevice, st));
        if (nvac > 0)
            TBG_HIP(hipMemcpyAsync(dvac.p, vac, (size_t)nvac * sizeof(int64_t),
                                   hipMemcpyHostToDevice, st));
        hipLaunchKernelGGL(k_hop_gen_count, dim3(blocks(norb_sc)), dim3(256), 0, st,
                           dorb.as<int>(), norb_sc, dind.as<int>(), nhf, g,
                           dvac.as<int64_t>(), nvac, cnt.as<int64_t>());
        TBG_HIP(hipGetLastError());
        size_t tmp_bytes = 0;
        TBG_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp_bytes, cnt.as<int64_t>(),
                                                 off.as<int64_t>(), norb_sc, st));
        DevBuf tmp(std::max<size_t>(tmp_bytes, 1));
        TBG_HIP(hipcub::DeviceScan::ExclusiveSum(tmp.p, tmp_bytes, cnt.as<int64_t>(),
                                                 off.as<int64_t>(), norb_sc, st));
        int64_t last_off = 0, last_cnt = 0;
        TBG_HIP(hipMemcpyAsync(&last_off, off.as<int64_t>() + norb_sc - 1, 8,
                               hipMemcpyDeviceToHost, st));
        TBG_HIP(hipMemcpyAsync(&last_cnt, cnt.as<int64_t>() + norb_sc - 1, 8,
                               hipMemcpyDeviceToHost, st));
        TBG_HIP(hipStreamSynchronize(st));
        const int64_t nfree = last_off + last_cnt;
        TBG_REQUIRE(total + nfree <= capacity, "output capacity too small");
        hipLaunchKernelGGL(k_hop_gen_fill, dim3(blocks(norb_sc)), dim3(256), 0, st,
                           dorb.as<int>(), norb_sc, dind.as<int>(), deng.as<double2>(), nhf, g,
                           dvac.as<int64_t>(), nvac, off.as<int64_t>(), dpos.as<double>(), total,
                           dhi.as<int64_t>(), dhj.as<int64_t>(), dhv.as<double2>(),
                           ddr.as<double>());
        TBG_HIP(hipGetLastError());
        total += nfree;
    }
    if (total > 0) {
        TBG_HIP(hipMemcpyAsync(hop_i, dhi.p, total * 8, hipMemcpyDeviceToHost, st));
        TBG_HIP(hipMemcpyAsync(hop_j, dhj.p, total * 8, hipMemcpyDeviceToHost, st));
        TBG_HIP(hipMemcpyAsync(hop_v, dhv.p, total * 16, hipMemcpyDeviceToHost, st));
        TBG_HIP(hipMemcpyAsync(dr, ddr.p, total * 24, hipMemcpyDeviceToHost, st));
    }
    TBG_HIP(hipStreamSynchronize(st));
    return total;
}

}  // namespace tbg
