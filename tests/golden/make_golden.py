"""Write the golden fixtures under tests/golden/ (run where the reference
sources are present: it needs oracle/_ref built from them).

Every fixture is data: inputs and the reference's outputs.
  bessel.npz     J_k series (scipy.special.jv, solver.py:194-224) for
                 tau = pi at 1e-14 and 1e-15 and tau = pi/2 at 1e-14
  fermi.npz      Fermi / one-minus-Fermi Chebyshev coefficients from the
                 reference's get_Fermi_cheb_coef (funcs.f90:60-123)
  graphene4.npz  4x4 periodic graphene: hop list, canonical CSR (scipy,
                 sample.py:375-391) and insertion-order CSR (sample.pyx:148-257)
  ref_dos.npz    12x12 graphene: initial state from the reference's
                 random_state (seed 1337) and C(t) from the reference tbpm_dos
  ref_ac.npz     8x8 graphene: states and C_ac(t) from the reference
                 tbpm_accond (2 samples)
  ref_dc.npz     8x8 graphene: states and the reference tbpm_dccond outputs
"""
import math
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402

KB = 8.617333262e-5


def graphene(L, rescale=9.0):
    hi, hj, hv, dr, pos = oracle.graphene_hops_loop(L, L)
    n = 2 * L * L
    orb = np.zeros(n) * (1.0 / rescale)
    hv = hv * (1.0 / rescale)
    return orb, hi, hj, hv, dr, pos


def main():
    if not oracle.ref_available():
        raise SystemExit("oracle/_ref/libtbpm_ref.so missing (make -C oracle ref)")
    pi = math.pi
    tau = 2 * pi / 18.0 * 9.0
    np.savez(os.path.join(HERE, "bessel.npz"),
             tau_pi=tau, tau_half=0.5 * tau,
             pi_1e14=oracle.bessel_series(tau, 1e-14),
             pi_1e15=oracle.bessel_series(tau, 1e-15),
             half_1e14=oracle.bessel_series(0.5 * tau, 1e-14))

    beta_re = 1.0 / (KB * 300) * 9.0
    np.savez(os.path.join(HERE, "fermi.npz"), beta_re=beta_re, mu_re=0.0,
             nr_fermi=2 ** 15, eps=1e-12,
             coef_F=oracle.ref_fermi_coef(2 ** 15, beta_re, 0.0, False, 1e-12),
             coef_omF=oracle.ref_fermi_coef(2 ** 15, beta_re, 0.0, True, 1e-12),
             beta_re_cold=beta_re * 3, mu_re_cold=0.1,
             coef_F_cold=oracle.ref_fermi_coef(2 ** 15, beta_re * 3, 0.1, False, 1e-12))

    orb, hi, hj, hv, dr, pos = graphene(4)
    ip, ix, dt = oracle.build_ham_csr(orb, hi, hj, hv)
    dip, dix, dh, ddx, ddy = oracle.build_ham_dxy_loop(orb, hi, hj, hv, dr)
    np.savez(os.path.join(HERE, "graphene4.npz"), hop_i=hi, hop_j=hj, hop_v=hv, dr=dr,
             orb_pos=pos, orb_eng=orb, csr_indptr=ip, csr_indices=ix, csr_data=dt,
             dxy_indptr=dip, dxy_indices=dix, dxy_hop=dh, dxy_dx=ddx, dxy_dy=ddy)

    orb, hi, hj, hv, dr, pos = graphene(12)
    n = len(orb)
    ip, ix, dt = oracle.build_ham_csr(orb, hi, hj, hv)
    bes = oracle.bessel_series(tau, 1e-15)
    T = 64
    psi = oracle.ref_random_state(n, 1337)
    corr = oracle.ref_tbpm_dos(bes, ip, ix, dt, 1337, T, 1)
    np.savez(os.path.join(HERE, "ref_dos.npz"), L=12, indptr=ip, indices=ix, data=dt, bes=bes,
             psi0=psi, corr=corr, T=T)

    orb, hi, hj, hv, dr, pos = graphene(8)
    n = len(orb)
    dip, dix, dh, ddx, ddy = oracle.build_ham_dxy_loop(orb, hi, hj, hv, dr)
    bes_h = oracle.bessel_series(0.5 * tau, 1e-14)
    T = 16
    S = 2
    seed = 1337
    psi = np.array([oracle.ref_random_state(n, seed * (s + 1)) for s in range(S)])
    ac = oracle.ref_tbpm_accond(bes_h, beta_re, 0.0, dip, dix, dh, 9.0, ddx, ddy, seed, T, S)
    np.savez(os.path.join(HERE, "ref_ac.npz"), L=8, indptr=dip, indices=dix, hop=dh, dx=ddx,
             dy=ddy, bes=bes_h, beta_re=beta_re, mu_re=0.0, rescale=9.0, psi0=psi, corr=ac, T=T)

    bes = oracle.bessel_series(tau, 1e-14)
    T = 16
    en_range = 18.0
    energies = np.array([0.5 * i * en_range / T - en_range / 2. for i in range(2 * T)])
    en_inds = np.where((energies >= -1.2) & (energies <= 1.2))[0]
    t_step = 2 * pi / en_range
    dos, dc = oracle.ref_tbpm_dccond(bes, beta_re, 0.0, dip, dix, dh, 9.0, ddx, ddy, seed, T, S,
                                     t_step, energies, en_inds)
    np.savez(os.path.join(HERE, "ref_dc.npz"), L=8, indptr=dip, indices=dix, hop=dh, dx=ddx,
             dy=ddy, bes=bes, rescale=9.0, psi0=psi, energies=energies, en_inds=en_inds,
             t_step=t_step, dos=dos, dc=dc, T=T)
    print("golden fixtures written")


if __name__ == "__main__":
    main()
