/*
 * tbplas_amd -- C ABI of the MI355X TBPM hot path.
 *
 * Plain pointers and sizes only.  Unless stated otherwise every array
 * argument is a DEVICE pointer on the handle's device; scalars and the small
 * coefficient tables (Bessel, Fermi, energies) are HOST pointers.
 * Complex numbers are interleaved (re, im) doubles, i.e. numpy complex128.
 * Every call is ordered on the handle's stream and returns after that stream
 * has finished the call's work.
 *
 * Status codes: 0 ok; TBG_ENORM the wave-function norm check failed (the
 * step is written to *fail_step, the outputs hold every step computed);
 * TBG_EINVAL bad argument; TBG_EOOM device allocation failed; TBG_EHIP a HIP
 * runtime call failed.  The message is available from tbg_last_error().
 * No call ever terminates the process (the reference STOPs,
 * tbplas/fortran/funcs.f90:175-192).
 */
#ifndef TBPLAS_AMD_H
#define TBPLAS_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TBG_OK 0
#define TBG_ENORM 1
#define TBG_EINVAL 2
#define TBG_EOOM 3
#define TBG_EHIP 4

#define TBG_FP64 0 /* complex128 storage and arithmetic */
#define TBG_FP32 1 /* complex64 storage and arithmetic, FP64 correlations */

typedef struct tbg_handle tbg_handle;
typedef struct tbg_ham tbg_ham;

/* ---- library / handle --------------------------------------------------- */
int tbg_version(void);
/* Message of the last failed call on this thread ("" if none). */
const char *tbg_last_error(void);
/* Create a handle on `device` with its own default-priority stream. */
int tbg_create(tbg_handle **h, int device);
int tbg_destroy(tbg_handle *h);
/* 1: replay each time step from a captured hipGraph (default); 0: launch. */
int tbg_set_graph(tbg_handle *h, int enable);
/* Launch shape of the propagation kernels: persistent blocks per compute
 * unit (1..64, default 16) and LDS staging of each slice's own rows (default
 * 1).  No reference counterpart; set by the tuner (tbplas_amd/tune.py). */
int tbg_set_tuning(tbg_handle *h, int blocks_per_cu, int stage);
/* Value layout of the Hamiltonians created on this handle from now on: with
 * enable = 1 (the default) a Hamiltonian whose hopping terms all have an
 * exactly zero imaginary part stores real values, which halves the value
 * bytes the propagation reads and gives the same sums.  No reference
 * counterpart (the reference always stores complex128,
 * tbplas/fortran/csr.f90:9-14). */
int tbg_set_real_layout(tbg_handle *h, int enable);
/* Several Chebyshev terms per pass for Hamiltonians created with spatial
 * tiles (tbg_ham_create_ordered): enable = 1 or 3 three terms, 4 four terms,
 * 5 five terms (each term more needs one more halo ring, so smaller tiles
 * fit the LDS), 0 one term per pass (same states, for comparison).  max_rows > 0 caps the rows one
 * workgroup owns (tiles are merged up to it; 0 = sized from the LDS).
 * Default (3, 0).  No reference counterpart. */
int tbg_set_blocking(tbg_handle *h, int enable, int max_rows);

/* ---- CSR assembly --------------------------------------------------------
 * Replaces the Cython build_ham_dxy_fast (tbplas/cython/sample.pyx:148-257,
 * called from tbplas/builder/sample.py:452-458).  Insertion order: within a
 * row, the kept diagonal (|orb_eng| >= cutoff) first, then for each hop ih
 * in order the entry (hop_i, hop_j, hop_v, +dr) and (hop_j, hop_i,
 * conj(hop_v), -dr).  Bit-exact in indptr/indices/hop/dx/dy.
 * Two phases: count writes indptr (n+1, device) and *nnz (host); the caller
 * allocates indices/hop/dx/dy with *nnz elements and calls fill.
 * dr is (m, 3) row-major float64; hop_v/hop are complex128.
 */
int tbg_ham_dxy_count(tbg_handle *h, int64_t n, int64_t m,
                      const double *orb_eng, const int64_t *hop_i,
                      const int64_t *hop_j, double cutoff, int64_t *indptr,
                      int64_t *nnz);
int tbg_ham_dxy_fill(tbg_handle *h, int64_t n, int64_t m,
                     const double *orb_eng, const int64_t *hop_i,
                     const int64_t *hop_j, const double *hop_v,
                     const double *dr, double cutoff, const int64_t *indptr,
                     int64_t *indices, double *hop, double *dx, double *dy);

/* Replaces Sample.build_ham_csr (tbplas/builder/sample.py:375-391), the
 * scipy sum dia(orb_eng) + csr(hop) + csr(hop)^H in canonical form: columns
 * sorted, duplicates summed ((dia + half) + half^H), exact zeros dropped.
 * Same two-phase protocol. */
int tbg_ham_csr_count(tbg_handle *h, int64_t n, int64_t m,
                      const double *orb_eng, const int64_t *hop_i,
                      const int64_t *hop_j, const double *hop_v,
                      int64_t *indptr, int64_t *nnz);
int tbg_ham_csr_fill(tbg_handle *h, int64_t n, int64_t m,
                     const double *orb_eng, const int64_t *hop_i,
                     const int64_t *hop_j, const double *hop_v,
                     const int64_t *indptr, int64_t *indices, double *data);

/* Replaces get_rescale (tbplas/cython/sample.pyx:75-114): the maximum over
 * rows of orb_eng[i] + sum |hop| of the hops touching row i.  Host result. */
int tbg_get_rescale(tbg_handle *h, int64_t n, int64_t m,
                    const double *orb_eng, const int64_t *hop_i,
                    const int64_t *hop_j, const double *hop_v,
                    double *rescale);

/* Graphene supercell arrays (HOST): replaces building make_graphene_diamond
 * (tbplas/materials/graphene.py:16-31) into a SuperCell and reading its
 * orbitals and hops (tbplas/builder/super.py:703-809,
 * tbplas/cython/super.pyx:383-986): same order and bit-identical values.
 * L1 x L2 cells, lattice constant a (nm), hopping t, pbc per direction,
 * vacancies (orbital ids before removal, may be NULL with nvac = 0).
 * The caller sizes orb_pos for 2*L1*L2 orbitals (x, y, z), hop_i/hop_j/dr
 * for 3*L1*L2 hops, hop_v as complex128; *n_out / *m_out get the counts. */
int tbg_graphene_hops(int64_t L1, int64_t L2, double t, int pbc1, int pbc2,
                      double a, const int64_t *vacancies, int64_t nvac,
                      double *orb_pos, int64_t *hop_i, int64_t *hop_j,
                      double *hop_v, double *dr, int64_t *n_out,
                      int64_t *m_out);

/* Replaces set_mag_field (tbplas/cython/sample.pyx:18-72): Peierls phases
 * for a field along -z, applied in place to hop_v.  HOST pointers (this is
 * O(m) host work done once per sample, kept on the host so the phases use
 * the same libm cos/sin as the reference).  gauge 0: (By, 0, 0);
 * 1: (0, -Bx, 0); 2: the half sum. */
int tbg_peierls_host(int64_t m, const int64_t *hop_i, const int64_t *hop_j,
                     double *hop_v, const double *dr, const double *orb_pos,
                     double intensity, int gauge);

/* ---- prepared Hamiltonian ------------------------------------------------
 * Stands in for make_csr_matrix (tbplas/fortran/csr.F90:45-68) and
 * current_coefficient (tbplas/fortran/funcs.f90:12-33): takes the CSR arrays
 * the reference passes to f2py (indptr/indices int64, hop complex128, all
 * device) and builds the device layout used by the propagation kernels.
 * dx/dy may be NULL; when given, the current operators
 * J_{x,y} = (0, cur_scale) * hop * d{x,y} are built on the same structure.
 * precision: TBG_FP64 or TBG_FP32.  Requires n < 2^31.
 */
int tbg_ham_create(tbg_handle *h, int64_t n, const int64_t *indptr,
                   const int64_t *indices, const double *hop,
                   const double *dx, const double *dy, double cur_scale,
                   int precision, tbg_ham **out);
/* tbg_ham_create with a row order for the device layout: perm (device,
 * int64, n, may be NULL) lists the caller's rows in the order the layout
 * stores them, so rows that are close in space can be placed in the same
 * slices (fewer far gathers).  tile_ptr (HOST, int64, ntiles + 1, may be
 * NULL) splits that order into spatial tiles (row offsets, 0 .. n); with
 * tiles, propagation runs three Chebyshev terms per pass over groups of
 * tiles plus their two halo rings (intermediate terms kept in LDS), which
 * reads the state vectors from HBM fewer times.  Inputs and outputs of every
 * entry point stay in the caller's order; each row keeps its entries'
 * order, so the states are bit-identical to the plain layout's.  No
 * reference counterpart (the reference propagates in the builder's order
 * one term per pass, tbplas/fortran/propagation.f90:42-48). */
int tbg_ham_create_ordered(tbg_handle *h, int64_t n, const int64_t *indptr,
                           const int64_t *indices, const double *hop,
                           const double *dx, const double *dy,
                           double cur_scale, int precision,
                           const int64_t *perm, const int64_t *tile_ptr,
                           int64_t ntiles, tbg_ham **out);
int tbg_ham_destroy(tbg_ham *H);
/* Stored elements (with slice padding) and maximum row width. */
int tbg_ham_info(const tbg_ham *H, int64_t *n, int64_t *nnz,
                 int64_t *padded, int *max_width);
/* *real = 1 when H stores real values (see tbg_set_real_layout). */
int tbg_ham_is_real(const tbg_ham *H, int *real);
/* Temporal blocking of H for `states` vectors of the given precision and
 * the handle's terms per pass (building its tables if needed): *nblocks
 * workgroup tiles (0 when H runs one term per pass) and *halo = rows
 * computed per row owned in the first term of a pass.  No reference
 * counterpart. */
int tbg_ham_blocking_info(tbg_handle *h, const tbg_ham *H, int states,
                          int precision, int *nblocks, double *halo);
/* Launch shape of the tile-pass kernel for the same arguments: *max_rows =
 * rows of the widest workgroup tile (halo included), *threads = threads
 * per workgroup (both 0 when H runs one term per pass).  No reference
 * counterpart. */
int tbg_ham_blocking_launch(tbg_handle *h, const tbg_ham *H, int states,
                            int precision, int *max_rows, int *threads);
/* What one full tile pass (all of the tables' terms, mode "continue") must
 * move, summed over its workgroup tiles: parts[0] = table bytes (columns,
 * values, LDS positions, row numbers, tile records), parts[1] = rows of
 * v_{m-1} gathered (distinct rows per tile), parts[2] = rows of v_{m-2}
 * read (R and its rings), parts[3] = owned rows (output read and written,
 * last two terms written: four row streams).  A row of states is
 * slots * element bytes.  All zero when H runs one term per pass.  No
 * reference counterpart. */
int tbg_ham_blocking_bytes(tbg_handle *h, const tbg_ham *H, int states,
                           int precision, int64_t *parts);

/* ---- per-launch timing -----------------------------------------------------
 * No reference counterpart.  With timing on, every propagation and
 * correlation launch of the handle is bracketed by HIP events on its stream
 * (graph replay is skipped meanwhile).  tbg_kernel_times synchronises the
 * stream, sums the durations per launch kind and clears them: kinds[i]
 * (1 = one-term kernel, 2 = AC correlation, 3 = DC quasi-eigenstate update,
 * 4 = DC current correlation, 5 = Hall moments' Gram block (k_gram),
 * 100 + 10 * mode + terms = tile pass with mode
 * 0 continue, 1 first terms from v_0, 2 last terms, 3 last terms with the
 * norm, 4 last terms with dot and norm), total_ms[i], counts[i], for at
 * most max_kinds kinds; *n_kinds = kinds seen. */
int tbg_set_kernel_timing(tbg_handle *h, int enable);
int tbg_kernel_times(tbg_handle *h, int max_kinds, int *kinds,
                     double *total_ms, int64_t *counts, int *n_kinds);

/* ---- correlation functions ----------------------------------------------
 * psi0: K initial states, complex128 (K, n) row-major (state k contiguous),
 * supplied by the caller (the reference draws them inside Fortran,
 * tbplas/fortran/random.f90:36-69).  Outputs are PER STATE; averaging over
 * states (the reference's corr += corrval / n_ran_samples) is the caller's.
 * bes: host Bessel series J_0..J_{n_bes-1} of the scaled time step
 * (tbplas/tbpm/solver.py:194-224), n_bes >= 3.
 * chk_step/chk_thr: norm check as tbplas/fortran/funcs.f90:175-192 at the
 * step the reference checks (chk_step outside the loop range disables it).
 */

/* Replaces f2py tbpm_dos (tbplas/fortran/tbpm.f90:6-93; called from
 * tbplas/tbpm/solver.py:271-281).  corr: complex128 (K, T+1), t = 0..T. */
int tbg_corr_dos(tbg_handle *h, const tbg_ham *H, const double *bes,
                 int n_bes, const double *psi0, int K, int T, int chk_step,
                 double chk_thr, double *corr, int *fail_step);

/* Replaces f2py tbpm_ldos (tbplas/fortran/tbpm.f90:97-189; called from
 * tbplas/tbpm/solver.py:286-315).  psi0: the LDOS states (random states
 * kept on the selected sites and normalised, tbpm.f90:151-156, formed by
 * the caller).  corr: complex128 (K, T+1) = <psi0|psi(t)>, t = 0..T.  The
 * norm check compares with the initial norm (tbpm.f90:170), chk_step in
 * 1..T. */
int tbg_corr_ldos(tbg_handle *h, const tbg_ham *H, const double *bes,
                  int n_bes, const double *psi0, int K, int T, int chk_step,
                  double chk_thr, double *corr, int *fail_step);

/* Replaces f2py tbpm_psi_t (tbplas/fortran/tbpm.f90:1029-1093; called from
 * tbplas/tbpm/solver.py:543-600).  psi0: device complex128 (n,); time_log:
 * host, n_log strictly increasing steps in [0, T]; psi_t: device
 * complex128 (n_log, n), row i = the state after time_log[i] steps.  The
 * norm of the state entering step chk_step is checked against psi0's. */
int tbg_psi_t(tbg_handle *h, const tbg_ham *H, const double *bes, int n_bes,
              const double *psi0, const int64_t *time_log, int n_log, int T,
              int chk_step, double chk_thr, double *psi_t, int *fail_step);

/* Replaces f2py tbpm_accond (tbplas/fortran/tbpm.f90:193-351; called from
 * tbplas/tbpm/solver.py:335-349).  H must carry current operators built with
 * cur_scale = rescale.  coef_F / coef_omF: host Chebyshev coefficients of the
 * Fermi and one-minus-Fermi operators (tbplas/fortran/funcs.f90:60-123).
 * corr: complex128 (K, 4, T) with columns xx, xy, yx, yy; column t=1 is zero
 * as in the reference (tbpm.f90:283-302 never stores it). */
int tbg_corr_ac(tbg_handle *h, const tbg_ham *H, const double *bes, int n_bes,
                const double *coef_F, int n_F, const double *coef_omF,
                int n_omF, const double *psi0, int K, int T, int chk_step,
                double chk_thr, double *corr, int *fail_step);

/* Replaces f2py tbpm_dccond (tbplas/fortran/tbpm.f90:497-718; called from
 * tbplas/tbpm/solver.py:422-436).  energies: host grid; en_inds: host 0-based
 * indices of the n_en quasi-eigenstate energies.  dos_corr: complex128
 * (K, T), t = 1..T.  dc_corr: complex128 (K, 2, n_en, T). */
int tbg_corr_dc(tbg_handle *h, const tbg_ham *H, const double *bes, int n_bes,
                const double *psi0, int K, int T, double t_step,
                const double *energies, const int64_t *en_inds, int n_en,
                int chk_step, double chk_thr, double *dos_corr,
                double *dc_corr, int *fail_step);

/* Replaces f2py tbpm_dyn_pol (tbplas/fortran/tbpm.f90:355-493; called from
 * tbplas/tbpm/solver.py:373-392).  coef_F / coef_omF: host Fermi and
 * one-minus-Fermi Chebyshev coefficients (as for tbg_corr_ac); q_points:
 * host (n_q, 3); orb_pos: host (n, 3) orbital positions in the caller's
 * order (density e^{i q.r}, funcs.f90:125-149).  corr: device float64
 * (K, n_q, T) = Im <F psi0 | e^{iq.r} U^{t-1} (1-F) e^{-iq.r} psi0> per state
 * (U^{t-1} acting on both), the reference's corr is the mean over states.
 * At most 32 states per call. */
int tbg_corr_dyn_pol(tbg_handle *h, const tbg_ham *H, const double *bes,
                     int n_bes, const double *coef_F, int n_F,
                     const double *coef_omF, int n_omF, const double *psi0,
                     int K, int T, const double *q_points, int n_q,
                     const double *orb_pos, int chk_step, double chk_thr,
                     double *corr, int *fail_step);

/* Replaces the Cython builders build_hop_pbc and build_hop_gen
 * (tbplas/cython/super.pyx:643-783, 860-986) as SuperCell.get_hop combines
 * them without a hop modifier (tbplas/builder/super.py:703-775): without
 * vacancies the periodic terms (in (cell, hop) order) then the terms that
 * cross an open boundary (in (orbital, hop) order); with vacancies every
 * term in (orbital, hop) order, dropping terms that leave an open boundary
 * or land on a vacancy.  All arrays are HOST arrays, the work runs on h's
 * device: pc_hop_ind (num_hop_pc, 5) int32 (ra, rb, rc, orb_i, orb_j),
 * pc_hop_eng complex128, dim/pbc (3,) int32, orb_id_pc (num_orb_sc, 4)
 * int32 (ia, ib, ic, orb) of the kept orbitals, vac_id_sc sorted vacancy
 * ids in the full-supercell numbering, sc_lattice (3, 3) Cartesian
 * supercell vectors, sc_orb_pos (num_orb_sc, 3).  Outputs of capacity
 * terms; *num_hop = the number written.  Bit-identical to the Cython. */
int tbg_supercell_hops(tbg_handle *h, const int32_t *pc_hop_ind,
                       const double *pc_hop_eng, int num_hop_pc,
                       const int32_t *dim, const int32_t *pbc, int num_orb_pc,
                       const int32_t *orb_id_pc, int64_t num_orb_sc,
                       const int64_t *vac_id_sc, int64_t num_vac,
                       const double *sc_lattice, const double *sc_orb_pos,
                       int64_t capacity, int64_t *hop_i, int64_t *hop_j,
                       double *hop_v, double *dr, int64_t *num_hop);

/* Haydock recursion coefficients of ldos_haydock (tbplas/fortran/
 * tbpm.f90:964-1024; Haydock_coef, propagation.f90:85-155; called from
 * tbplas/tbpm/solver.py:507-541) for K starting states psi0 (device
 * complex128 (K, n), normalised, non-zero on the selected sites), H ->
 * rescale * H~ (amv).  coefa: HOST complex128 (K, depth) = a_i, coefb: HOST
 * float64 (K, depth) = b_i; the Green's function and the LDOS from their
 * averages stay with the caller (green_function, funcs.f90:151-172). */
int tbg_haydock_coef(tbg_handle *h, const tbg_ham *H, const double *psi0,
                     int K, int depth, double rescale, double *coefa,
                     double *coefb);

/* Replaces f2py tbpm_kbdc (tbplas/fortran/tbpm.f90:836-960; called from
 * tbplas/tbpm/solver.py:459-468).  H: complex128 with current operators
 * (cur_scale = rescale).  psi0: device complex128 (K, n).  mu: device
 * complex128 (K, n_kernel, n_kernel), state s in column-major order like
 * the reference's corr_mu (element (i, j) at s*N*N + i + j*N), already
 * multiplied by the Jackson weights g_i g_j (kpm.f90:8-29); the
 * reference's corr_mu_avg is the mean over states.  direction 1: xx, 2:
 * xy. */
int tbg_hall_mu(tbg_handle *h, const tbg_ham *H, const double *psi0, int K,
                int n_kernel, int direction, double *mu);

/* Replaces f2py tbpm_eigenstates (tbplas/fortran/tbpm.f90:721-830; called
 * from tbplas/tbpm/solver.py:472-505).  energies: host, n_en energies (eV,
 * with t_step in hbar/eV as for tbg_corr_dc).  qe: device float64
 * (K, n_en, n) = |QE|^2 per state after normalisation (tbpm.f90:812-825);
 * the reference's wf_QE is the average over states of these rows.  At most
 * 32 states per call (each propagates forward and backward). */
int tbg_quasi_eigenstates(tbg_handle *h, const tbg_ham *H, const double *bes,
                          int n_bes, const double *psi0, int K, int T,
                          double t_step, const double *energies, int n_en,
                          int chk_step, double chk_thr, double *qe,
                          int *fail_step);

/* One propagation step psi <- exp(-+ i H~ tau) psi of K states in place
 * (cheb_wf_timestep, tbplas/fortran/propagation.f90:12-49).  fwd=1: forward.
 * psi: complex128 (K, n). */
int tbg_cheb_step(tbg_handle *h, const tbg_ham *H, const double *bes,
                  int n_bes, double *psi, int K, int fwd);

/* Apply a real Chebyshev polynomial sum_k c_k T_k(H~) to K states in place
 * (Fermi, tbplas/fortran/propagation.f90:52-82).  psi: complex128 (K, n). */
int tbg_cheb_poly(tbg_handle *h, const tbg_ham *H, const double *coef,
                  int n_coef, double *psi, int K);

/* ---- batched uneven instances --------------------------------------------
 * No reference counterpart (the reference runs one sample at a time): DOS
 * correlations of several independent Hamiltonians in one launch per
 * Chebyshev term.  hams: n_inst prepared Hamiltonians (same precision);
 * psi0[i]: device complex128 (K, n_i); corr[i]: device complex128 (K, T+1).
 */
int tbg_corr_dos_batch(tbg_handle *h, int n_inst, const tbg_ham *const *hams,
                       const double *bes, int n_bes,
                       const double *const *psi0, int K, int T, int chk_step,
                       double chk_thr, double *const *corr, int *fail_inst,
                       int *fail_step);

/* tbg_corr_dos without device allocation.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same arguments, outputs, norm check and status codes as tbg_corr_dos; the
 * run's buffers (recurrence vectors, packed states, partial sums,
 * correlation rows) come from a 64 MiB device array the library holds on
 * every GPU it runs on until the process ends, so the run must fit it:
 * about 4 * n_pad * slots * (8 or 16) bytes + (T + 1) * 24 * slots bytes
 * (n_pad = n rounded up to 64: about one million orbitals with one
 * complex128 state, 260,000 with 8 complex64 states);
 * larger runs return TBG_EINVAL.  One term per pass (spatial tiles are not
 * used), so the correlations equal tbg_corr_dos's bit for bit on a
 * Hamiltonian without tiles.  Every call on a GPU shares its array and
 * leaves its contents there: calls that overlap in time on one GPU race. */
int tbg_corr_dos_noalloc(tbg_handle *h, const tbg_ham *H, const double *bes,
                         int n_bes, const double *psi0, int K, int T,
                         int chk_step, double chk_thr, double *corr,
                         int *fail_step);

/* ---- DOS run in pieces ----------------------------------------------------
 * tbg_corr_dos / tbg_corr_dos_batch split into set-up, propagation and
 * read-out, so a caller can time the propagation alone or stop early:
 * begin packs the states and forms C(0); advance runs nsteps more steps
 * (at most T in total) and reports their device time in milliseconds (HIP
 * events on the handle's stream); end writes corr for the steps run so far
 * (later steps stay zero) and applies the norm check when chk_step was
 * reached.  The run must be destroyed with tbg_run_destroy. */
typedef struct tbg_run tbg_run;
int tbg_dos_begin(tbg_handle *h, int n_inst, const tbg_ham *const *hams,
                  const double *bes, int n_bes, const double *const *psi0,
                  int K, int T, tbg_run **out);
int tbg_run_advance(tbg_run *r, int nsteps, double *elapsed_ms);
int tbg_dos_end(tbg_run *r, int chk_step, double chk_thr,
                double *const *corr, int *fail_inst, int *fail_step);
int tbg_run_destroy(tbg_run *r);
/* The byte counts of tbg_ham_blocking_bytes for the tables the run uses
 * (a batch's concatenated tiles included) and the bytes of one row of
 * states; all zero when the run goes one term per pass. */
int tbg_run_blocking_bytes(const tbg_run *r, int64_t *parts, int *row_bytes);

/* ---- learned DOS surrogate -------------------------------------------------
 * No reference counterpart.  A dense network x -> GELU(x W_0 + b_0) -> ... ->
 * x W_{L-1} + b_{L-1} on the matrix cores: BF16 operands, FP32 accumulation,
 * FP32 master weights, Adam.  widths: n_layers+1 entries (input, hidden...,
 * output).  Parameters are packed per layer as W_l [in][out] row-major, then
 * b_l [out]; set/get use HOST FP32 arrays.  x, target, y are DEVICE FP32
 * row-major [B][features].  The loss is the mean squared error. */
typedef struct tbg_mlp tbg_mlp;
int tbg_mlp_create(tbg_handle *h, int n_layers, const int *widths,
                   tbg_mlp **out);
int tbg_mlp_destroy(tbg_mlp *m);
int tbg_mlp_num_params(const tbg_mlp *m, int64_t *n);
int tbg_mlp_set_params(tbg_mlp *m, const float *params, int64_t n);
int tbg_mlp_get_params(const tbg_mlp *m, float *params, int64_t n);
int tbg_mlp_forward(tbg_mlp *m, const float *x, int B, float *y);
/* loss and host gradients (packed like the parameters), no update */
int tbg_mlp_gradients(tbg_mlp *m, const float *x, const float *target, int B,
                      float *grads, double *loss);
/* `steps` Adam steps on one batch; *loss = loss before the last step */
int tbg_mlp_train(tbg_mlp *m, const float *x, const float *target, int B,
                  int steps, float lr, float beta1, float beta2, float eps,
                  double *loss);
/* C [M][N] (FP32) = A [M][K] B [K][N], row-major BF16 bit patterns, on the
 * same MFMA kernel (device pointers) */
int tbg_gemm_bf16(tbg_handle *h, int M, int N, int K, const uint16_t *A,
                  const uint16_t *B, float *C);

#ifdef __cplusplus
}
#endif
#endif /* TBPLAS_AMD_H */
