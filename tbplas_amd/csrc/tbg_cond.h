// Conductivity drivers (AC, DC) and batch concatenation (tbg_cond.hip).
#pragma once

#include "tbg_engine.h"

namespace tbg {

// Block-diagonal concatenation of prepared Hamiltonians.
struct Concat {
    DevBuf cols, vals, soff, sw;
    int64_t nslices = 0;
    int real_vals = 0;
    MatView view() const {
        MatView m;
        m.cols = cols.as<int>();
        m.vals = vals.p;
        m.soff = soff.as<int64_t>();
        m.sw = sw.as<int>();
        m.nslices = nslices;
        m.real_vals = real_vals;
        return m;
    }
};
void concat_hams(tbg_handle* h, int n, const tbg_ham* const* hams, Concat& out);

template <typename R>
void corr_ac_impl(tbg_handle* h, const tbg_ham* H, const double* bes, int nb, const double* cF,
                  int nF, const double* cO, int nO, const double* psi0, int K, int T,
                  int chk_step, double chk_thr, double* corr, int* fail_step);

// tbg_hall.hip (coefa, coefb: host outputs)
void haydock_impl(tbg_handle* h, const tbg_ham* H, const double* psi0, int K, int depth,
                  double rescale, double* coefa, double* coefb);
void hall_mu_impl(tbg_handle* h, const tbg_ham* H, const double* psi0, int K, int N, int dir,
                  double* mu);

template <typename R>
void dyn_pol_impl(tbg_handle* h, const tbg_ham* H, const double* bes, int nb, const double* cF,
                  int nF, const double* cO, int nO, const double* psi0, int K, int T,
                  const double* q_points, int n_q, const double* pos, int chk_step,
                  double chk_thr, double* corr, int* fail_step);

template <typename R>
void corr_dc_impl(tbg_handle* h, const tbg_ham* H, const double* bes, int nb, const double* psi0,
                  int K, int T, double t_step, const double* energies, const int64_t* en_inds,
                  int n_en, int chk_step, double chk_thr, double* dos_corr, double* dc_corr,
                  int* fail_step, double* qe_abs2 = nullptr);

}  // namespace tbg
