"""The benchmark configurations at their own sizes and layouts, against the
oracle.

* 2048x2048 graphene (8,388,608 orbitals), 8 states, 2048 steps, five
  Chebyshev terms per pass on the tiles the benchmark times (complex64
  4.8 nm, complex128 3.2 nm) and four on the edges sized for four
  (complex64 5.0 nm, complex128 3.4 nm; complex128 also on 5.0 nm, split
  to fit): C(t) at sampled steps over the whole run against exact k-space
  propagation.  Tolerances: complex128 1e-10;
  complex64 (FP32 storage and arithmetic, FP64 sums) 1e-6, about ten times
  the largest error measured over 2048 steps.
* AC on the same sample (8 states in runs of 4 on the AC rule's tiles,
  complex64 3.8 nm and complex128 2.5 nm, with five terms per pass, on
  2.5 nm with four, and in one run on 1.5 nm tiles with three) against exact
  k-space propagation with the product's Fermi polynomial over 256 steps:
  complex128 1e-11, complex64 1e-6 (correlations ~4e-2).
* The disordered, field-threaded DC sample (1% vacancies, Anderson disorder
  U[-0.5, 0.5] eV, B = 20 T in the Landau gauge, 16 states, 5 energies,
  2.0 nm tiles with three, four and five terms per pass) on 24x24 cells against
  the C restatement of tbpm_dccond, and at 1024x1024 (2,076,180 orbitals)
  through properties that hold at any size: the DC run's DOS correlation
  equals a DOS run's C(t) on the same states, and the quasi-eigenstates are
  normalised.
"""
import math

import numpy as np
import pytest

import oracle
import tbplas_amd as tb
from tbplas_amd import tbpm
from conftest import graphene_sample, tau_of

pytestmark = pytest.mark.gpu

L2048 = 2048
STEPS = 2048
# sampled steps: the first few, every 128th and the last
PTS = sorted(set([0, 1, 2, 3, 4, 5, 17, 21, 100, 511, 1023, 1025, 2047, STEPS]
                 + list(range(0, STEPS + 1, 128))))


def _states(n, K, seed=1337):
    return tb.states.random_states(n, tb.states.sample_seeds(seed, K))


@pytest.fixture(scope="module")
def graphene2048():
    s = graphene_sample(L2048)
    ip, ix, dt = s.build_ham_csr_device()
    bes = tbpm.bessel_series(tau_of(s), 1e-15)
    psi = _states(s.num_orb, 8)
    exact = oracle.kspace_graphene_corr_multi(L2048, L2048, s.hop_v[0], psi, tau_of(s), PTS)
    return s, (ip, ix, dt), bes, psi, exact


@pytest.mark.parametrize("prec,tol,tile,terms", [("fp32", 1e-6, 4.8, 5), ("fp64", 1e-10, 3.2, 5),
                                                 ("fp32", 1e-6, 5.0, 4), ("fp64", 1e-10, 3.4, 4),
                                                 ("fp64", 1e-10, 5.0, 4)])
def test_benchmark_layout_2048_steps_vs_kspace(gpu, graphene2048, prec, tol, tile, terms):
    """The benchmark's kernel configurations over the whole 2048-step run,
    16-wave workgroups on the tile edge the benchmark takes for the leg
    (the library's rule to 0.1 nm) with five terms per pass (terms 1-5
    from v_0, two full passes, terms 16-20 with the correlation sums per
    tile): 8 complex64 states a row on 4.8 nm tiles, 8 complex128 states on
    3.2 nm tiles; with four terms (terms 1-4, three full passes, terms
    17-20) on the edges sized for four: complex64 on 5.0 nm, complex128 on
    3.4 nm; and complex128 on 5.0 nm tiles, which the library splits to fit
    the LDS, so that the output partials of owned rows beyond the register
    queue go through memory (two held per lane, up to eight owned rows per
    lane)."""
    s, (ip, ix, dt), bes, psi, exact = graphene2048
    assert len(bes) == 21
    assert round(tb.auto_tile_nm(8, "fp32"), 1) == 4.8
    assert round(tb.auto_tile_nm(8, "fp32", terms=4), 1) == 5.0
    assert round(tb.auto_tile_nm(8, "fp64"), 1) == 3.2
    assert round(tb.auto_tile_nm(8, "fp64", terms=4), 1) == 3.4
    order, tiles = s.tile_order(tile, return_tiles=True)
    h = tb._lib.Handle(0)
    try:
        h.set_tuning(8, True)
        h.set_blocking(True, terms=terms)
        H = tbpm.Hamiltonian(s.num_orb, ip, ix, dt, precision=prec, handle=h, order=order,
                             tiles=tiles)
        assert H.info()["real_values"]
        nb, halo = H.blocking_info(8)
        rows, nt = H.blocking_launch(8)
        assert nb > 4000
        assert nt == 1024
        if prec == "fp32" or tile < 4.0:
            # whole tiles fit the LDS (five terms reach two rings further)
            assert 1.0 < halo < (1.6 if terms == 4 else 1.75)
        else:
            # 128-byte rows on 5.0 nm: the tiles are split to fit
            assert 1.5 < halo < 2.0
        c = tbpm.corr_dos(H, bes, psi, STEPS, chk_thr=1e-4)
        H.close()
    finally:
        h.close()
    err = float(np.max(np.abs(c[:, PTS] - exact)))
    print(f"{prec} K=8 2048x2048, {tile} nm tiles, {terms} terms per pass, {nb} tiles, halo "
          f"{halo:.3f} "
          f"(widest {rows} rows, {nt} threads): max |C - C_exact| over {STEPS} steps = "
          f"{err:.3e}", flush=True)
    assert err < tol
    # C(0) is the squared norm of the state as stored (complex64 rounds it)
    assert np.max(np.abs(c[:, 0] - 1.0)) < (1e-9 if prec == "fp32" else 1e-12)


AC_STEPS = 256
AC_PTS = [2, 3, 4, 17, 64, 128, 255, AC_STEPS]


@pytest.fixture(scope="module")
def ac2048(graphene2048):
    """The AC configuration of the benchmark sample: 300 K, mu = 0, 8 states,
    tau = pi/2, the Fermi series of the product, and its exact k-space
    correlations at sampled steps of a 256-step run."""
    s, _, _, psi, _ = graphene2048
    ip, ix, hop, dx, dy = s.build_ham_dxy_device()
    tau = tau_of(s, 0.5)
    bes = tbpm.bessel_series(tau, 1e-14)
    beta = 1.0 / (8.617333262e-5 * 300.0) * s.rescale
    cF = tbpm.fermi_cheb_coef(2 ** 15, beta, 0.0, False, 1e-12)
    cO = tbpm.fermi_cheb_coef(2 ** 15, beta, 0.0, True, 1e-12)
    exact = oracle.kspace_graphene_ac(L2048, L2048, s.hop_v[:3], s.dr[:3], s.rescale, cF, cO,
                                      psi, tau, AC_PTS)
    return s, (ip, ix, hop, dx, dy), bes, (cF, cO), psi, exact


@pytest.mark.parametrize("prec,tol,tile,terms,per", [
    ("fp32", 1e-6, 3.8, 5, 4), ("fp32", 1e-6, 2.5, 4, 4), ("fp64", 1e-11, 2.5, 4, 4),
    ("fp64", 1e-11, 2.5, 5, 4), ("fp32", 1e-6, 1.5, 3, 8), ("fp64", 1e-11, 1.5, 3, 8)])
def test_ac_benchmark_layout_vs_kspace(gpu, ac2048, prec, tol, tile, terms, per):
    """AC on 8.4M-orbital graphene as the benchmark and the Solver run it
    (runs of 4 states, 12 vectors: J_x psi1, J_y psi1 and psi2 of each, on
    the tiles of the library's AC rule with five terms per pass, which the
    tuner picks there: complex64 3.8 nm, complex128 2.5 nm; both also on
    2.5 nm with four) and as one run of 8 states (24
    vectors, 1.5 nm tiles, three terms per pass), against exact k-space
    propagation with the same Fermi polynomial, at sampled steps of 256
    (the correlations are ~4e-2)."""
    s, (ip, ix, hop, dx, dy), bes, (cF, cO), psi, exact = ac2048
    order, tiles = s.tile_order(tile, return_tiles=True)
    h = tb._lib.Handle(0)
    try:
        h.set_blocking(True, terms=terms)
        H = tbpm.Hamiltonian(s.num_orb, ip, ix, hop, dx, dy, cur_scale=s.rescale, precision=prec,
                             handle=h, order=order, tiles=tiles)
        assert H.blocking_info(3 * per)[0] > 1000
        c = np.concatenate([tbpm.corr_ac(H, bes, cF, cO, psi[a:a + per], AC_STEPS, chk_thr=1e-4)
                            for a in range(0, len(psi), per)])
        H.close()
    finally:
        h.close()
    got = c[:, :, [t - 1 for t in AC_PTS]]
    err = float(np.max(np.abs(got - exact)))
    print(f"AC {prec} 2048x2048, 8 states in runs of {per}, {tile} nm tiles, {AC_STEPS} "
          f"steps: max |C - C_exact| = {err:.3e} "
          f"(max |C| {np.max(np.abs(exact)):.3e})", flush=True)
    assert np.max(np.abs(c[:, :, 0])) == 0.0
    assert err < tol
    if tile != 1.5 and per == 4:
        assert tile == 2.5 or tile == round(tb.ac_tile_nm(3 * per, prec), 1)


def _cfg3_sample(L, seed=2024):
    s = tb.make_graphene_sample(L, vacancies=tb.random_vacancies(2 * L * L, 0.01, seed))
    tb.add_anderson_disorder(s, 1.0, seed + 1)
    s.set_magnetic_field(20.0, 0)
    s.rescale_ham(9.0)
    return s


def _en_grid(s, T):
    en = np.array([0.5 * i * s.energy_range / T - s.energy_range / 2. for i in range(2 * T)])
    inds = np.array([int(np.argmin(np.abs(en - e))) for e in (-1.0, -0.5, 0.0, 0.5, 1.0)])
    return en, inds


@pytest.mark.parametrize("terms", [3, 4, 5])
def test_cfg3_dc_vs_oracle(gpu, terms):
    """1% vacancies, Anderson W = 1 eV, B = 20 T, 16 states (32 propagated
    vectors), 5 energies in [-1, 1] eV, 2.0 nm tiles: complex128 to 1e-10
    and complex64 to 2e-4 (relative to the largest correlation) against
    orc_tbpm_dccond over 40 steps."""
    s = _cfg3_sample(24)
    assert np.any(s.hop_v.imag != 0)
    ip, ix, hop, dx, dy = oracle.build_ham_dxy_loop(s.orb_eng, s.hop_i, s.hop_j, s.hop_v, s.dr)
    order, tiles = s.tile_order(2.0, return_tiles=True)
    T, K = 40, 16
    ts = 2 * math.pi / s.energy_range
    bes = tbpm.bessel_series(ts * s.rescale, 1e-14)
    en, inds = _en_grid(s, T)
    psi = _states(s.num_orb, K, seed=77)
    rd, rc = oracle.orc_tbpm_dccond(ip, ix, hop, dx, dy, s.rescale, bes, psi, T, ts, en, inds)
    h = tb._lib.Handle(0)
    try:
        h.set_blocking(True, terms=terms)
        for prec, tol in (("fp64", 1e-10), ("fp32", 2e-4)):
            H = tbpm.Hamiltonian(s.num_orb, ip, ix, hop, dx, dy, cur_scale=s.rescale,
                                 precision=prec, handle=h, order=order, tiles=tiles)
            assert not H.info()["real_values"]
            nb = H.blocking_info(2 * K)[0]
            # complex64: 32 vectors of 8 bytes a row fit tile passes at
            # 2.0 nm; complex128 may fall back to one term per pass
            assert nb > 1 or prec == "fp64"
            d, c = tbpm.corr_dc(H, bes, psi, T, ts, en, inds, chk_thr=1e-4)
            H.close()
            scale = max(1.0, float(np.max(np.abs(rc))))
            assert np.max(np.abs(d - rd)) < tol, prec
            assert np.max(np.abs(c - rc)) < tol * scale, prec
    finally:
        h.close()


def test_cfg3_full_size_properties(gpu):
    """1024x1024 with 1% vacancies (2,076,180 orbitals), Anderson W = 1 eV,
    B = 20 T, complex64, 2.0 nm tiles: the DOS correlation of the DC run
    (t = 1..T) equals the DOS run's C(t) on the same 16 states, and every
    quasi-eigenstate has unit norm."""
    s = _cfg3_sample(1024)
    assert s.num_orb == 2076180
    ip, ix, hop, dx, dy = s.build_ham_dxy_device()
    order, tiles = s.tile_order(2.0, return_tiles=True)
    T, K = 256, 16
    ts = 2 * math.pi / s.energy_range
    bes = tbpm.bessel_series(ts * s.rescale, 1e-14)
    en, inds = _en_grid(s, T)
    psi = _states(s.num_orb, K, seed=5)
    h = tb._lib.Handle(0)
    try:
        h.set_blocking(True, terms=4)
        H = tbpm.Hamiltonian(s.num_orb, ip, ix, hop, dx, dy, cur_scale=s.rescale,
                             precision="fp32", handle=h, order=order, tiles=tiles)
        d, c = tbpm.corr_dc(H, bes, psi, T, ts, en, inds, chk_thr=1e-4)
        assert np.all(np.isfinite(c))
        cd = tbpm.corr_dos(H, bes, psi, T, chk_thr=1e-4)
        err = float(np.max(np.abs(d - cd[:, 1:])))
        print(f"cfg3 full size: max |dos_corr(DC) - C_DOS(t)| = {err:.3e}", flush=True)
        assert err < 1e-6
        qe = tbpm.quasi_eigenstates(H, bes, psi[:2], 64, ts, en[inds], chk_thr=1e-4)
        H.close()
    finally:
        h.close()
    norms = qe.sum(axis=2)
    assert np.max(np.abs(norms - 1.0)) < 1e-10
