// CSR assembly and device-layout builders (tbg_csr.hip).
#pragma once

#include "tbg_internal.h"

namespace tbg {

void scan_i64(tbg_handle* h, const int64_t* in, int64_t n, int64_t* out, bool exclusive,
              DevBuf& tmp);

// Row pointer of the emission stream (diag_mode 1: |orb_eng| >= cutoff,
// 2: orb_eng != 0, 0: no diagonal).  Returns the entry count.
int64_t place_count(tbg_handle* h, int64_t n, int64_t m, const double* orb_eng, int diag_mode,
                    double cutoff, const int64_t* hop_i, const int64_t* hop_j, int64_t* indptr,
                    DevBuf& tmp);

void dxy_fill(tbg_handle* h, int64_t n, int64_t m, const double* orb_eng, const int64_t* hop_i,
              const int64_t* hop_j, const double* hop_v, const double* dr, double cutoff,
              const int64_t* indptr, int64_t* indices, double* hop, double* dx, double* dy);

int64_t canon_build(tbg_handle* h, int64_t n, int64_t m, const double* orb_eng,
                    const int64_t* hop_i, const int64_t* hop_j, const double* hop_v,
                    int64_t* indptr_out, const int64_t* indptr_in, int64_t* indices,
                    double* data);

double get_rescale(tbg_handle* h, int64_t n, int64_t m, const double* orb_eng,
                   const int64_t* hop_i, const int64_t* hop_j, const double* hop_v);

// perm (device, n, optional): new row -> caller row; the layout stores the
// rows in that order with renumbered columns.
void build_sell(tbg_handle* h, tbg_ham* H, const int64_t* indptr, const int64_t* indices,
                const double* hop, const double* dx, const double* dy, double cur_scale,
                const int64_t* perm);

// tbg_super.hip: SuperCell.get_hop without a hop modifier (host arrays in
// and out, work on h's device); returns the number of terms.
int64_t supercell_hops(tbg_handle* h, const int* pc_hop_ind, const double* pc_hop_eng, int nh,
                       const int* dim, const int* pbc, int norb_pc, const int* orb_id_pc,
                       int64_t norb_sc, const int64_t* vac, int64_t nvac, const double* sc_lat,
                       const double* sc_pos, int64_t capacity, int64_t* hop_i, int64_t* hop_j,
                       double* hop_v, double* dr);

}  // namespace tbg
