"""GPU TBPM correlation kernels against the oracle.

Tolerances: FP64 path 1e-12 on small samples and 1e-10 on the 131,072-orbital
1024-step case; FP32 path (complex64 storage, FP64 correlation sums)
2e-5 max |dC| over 64 steps and 2e-4 over 1024 steps.
"""
import math

import numpy as np
import pytest

import oracle
import tbplas_amd as tb
from tbplas_amd import tbpm
from conftest import golden, graphene_sample, tau_of

pytestmark = pytest.mark.gpu


def _csr(s):
    return oracle.build_ham_csr(s.orb_eng, s.hop_i, s.hop_j, s.hop_v)


def _states(n, K, seed=1337):
    return tb.states.random_states(n, tb.states.sample_seeds(seed, K))


def test_dos_reference_fixture_fp64(gpu):
    """Identical initial state to the reference Fortran run (its own RNG)."""
    g = golden("ref_dos.npz")
    n = len(g["indptr"]) - 1
    H = tbpm.Hamiltonian(n, g["indptr"], g["indices"], g["data"], precision="fp64")
    c = tbpm.corr_dos(H, g["bes"], g["psi0"][None], int(g["T"]))[0]
    assert np.max(np.abs(c - g["corr"])) < 1e-12


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-12), ("fp32", 2e-5)])
@pytest.mark.parametrize("K", [1, 2, 3, 8])
def test_dos_vs_oracle(gpu, prec, tol, K):
    s = graphene_sample(20)
    ip, ix, dt = _csr(s)
    bes = tbpm.bessel_series(tau_of(s), 1e-15)
    psi = _states(s.num_orb, K)
    T = 64
    ref = oracle.orc_tbpm_dos(ip, ix, dt, bes, psi, T)
    H = tbpm.Hamiltonian(s.num_orb, ip, ix, dt, precision=prec)
    c = tbpm.corr_dos(H, bes, psi, T)
    assert c.shape == (K, T + 1)
    assert np.max(np.abs(c - ref)) < tol


@pytest.mark.parametrize("nb_prec", [1e-14, 1e-15, 1e-6])
def test_dos_disordered_complex_hops(gpu, nb_prec):
    """Vacancies, on-site disorder (4-entry rows), Peierls phases; different
    series lengths exercise every accumulation schedule of the fused kernel."""
    s = tb.make_graphene_sample(18, vacancies=tb.random_vacancies(2 * 18 * 18, 0.02, 3))
    tb.add_anderson_disorder(s, 0.8, 4)
    s.set_magnetic_field(40.0, 0)
    s.rescale_ham(9.0)
    ip, ix, dt = _csr(s)
    bes = tbpm.bessel_series(tau_of(s), nb_prec)
    psi = _states(s.num_orb, 2, seed=5)
    T = 40
    ref = oracle.orc_tbpm_dos(ip, ix, dt, bes, psi, T)
    H = tbpm.Hamiltonian(s.num_orb, ip, ix, dt, precision="fp64")
    c = tbpm.corr_dos(H, bes, psi, T)
    assert np.max(np.abs(c - ref)) < 1e-12


@pytest.mark.parametrize("nterms", [3, 4, 5, 6, 7, 8])
def test_short_series_schedules(gpu, nterms):
    s = graphene_sample(8)
    ip, ix, dt = _csr(s)
    bes = np.asarray(tbpm.bessel_series(tau_of(s), 1e-15)[:nterms])
    psi = _states(s.num_orb, 1)
    ref = oracle.orc_tbpm_dos(ip, ix, dt, bes, psi, 6, chk_step=0)
    H = tbpm.Hamiltonian(s.num_orb, ip, ix, dt, precision="fp64")
    c = tbpm.corr_dos(H, bes, psi, 6, chk_step=0)
    assert np.max(np.abs(c - ref)) < 1e-13


def test_multivector_equals_single_runs(gpu):
    s = graphene_sample(16)
    ip, ix, dt = _csr(s)
    bes = tbpm.bessel_series(tau_of(s), 1e-15)
    psi = _states(s.num_orb, 8)
    H = tbpm.Hamiltonian(s.num_orb, ip, ix, dt, precision="fp64")
    many = tbpm.corr_dos(H, bes, psi, 32)
    for k in range(8):
        one = tbpm.corr_dos(H, bes, psi[k:k + 1], 32)[0]
        assert np.max(np.abs(one - many[k])) < 1e-14


def test_graph_replay_equals_launches(gpu):
    s = graphene_sample(16)
    ip, ix, dt = _csr(s)
    bes = tbpm.bessel_series(tau_of(s), 1e-15)
    psi = _states(s.num_orb, 4)
    H = tbpm.Hamiltonian(s.num_orb, ip, ix, dt, precision="fp32")
    H.handle.set_graph(True)
    a = tbpm.corr_dos(H, bes, psi, 70)      # 4 replays of 16 steps + 6 launched
    b = tbpm.corr_dos(H, bes, psi, 70)
    H.handle.set_graph(False)
    c = tbpm.corr_dos(H, bes, psi, 70)
    H.handle.set_graph(True)
    np.testing.assert_array_equal(a, b)
    np.testing.assert_array_equal(a, c)


@pytest.mark.parametrize("prec,K", [("fp32", 8), ("fp32", 1), ("fp64", 1), ("fp64", 4)])
def test_launch_shapes_agree(gpu, prec, K):
    """LDS staging on/off gives identical states; other block counts only
    reorder the FP64 correlation sums."""
    s = tb.make_graphene_sample(24, vacancies=tb.random_vacancies(2 * 24 * 24, 0.02, 1))
    tb.add_anderson_disorder(s, 0.5, 2)
    s.rescale_ham(9.0)
    ip, ix, dt = _csr(s)
    bes = tbpm.bessel_series(tau_of(s), 1e-15)
    psi = _states(s.num_orb, K)
    H = tbpm.Hamiltonian(s.num_orb, ip, ix, dt, precision=prec)
    h = H.handle
    try:
        h.set_tuning(8, True)
        a = tbpm.corr_dos(H, bes, psi, 20, chk_thr=1e-4)
        h.set_tuning(8, False)
        b = tbpm.corr_dos(H, bes, psi, 20, chk_thr=1e-4)
        np.testing.assert_array_equal(a, b)
        for bpc in (1, 3, 16):
            h.set_tuning(bpc, True)
            c = tbpm.corr_dos(H, bes, psi, 20, chk_thr=1e-4)
            assert np.max(np.abs(c - a)) < 1e-14
    finally:
        h.set_tuning(8, True)


@pytest.mark.parametrize("prec,K", [("fp32", 8), ("fp32", 1), ("fp64", 1), ("fp64", 4)])
def test_real_layout_equals_complex_layout(gpu, prec, K):
    """Real-valued storage of a Hamiltonian without imaginary parts skips
    only products with exact zeros: identical correlations."""
    s = tb.make_graphene_sample(24, vacancies=tb.random_vacancies(2 * 24 * 24, 0.02, 1))
    tb.add_anderson_disorder(s, 0.5, 2)
    s.rescale_ham(9.0)
    ip, ix, dt = _csr(s)
    assert not np.any(dt.imag)
    bes = tbpm.bessel_series(tau_of(s), 1e-15)
    psi = _states(s.num_orb, K)
    h = tb._lib.Handle(0)
    try:
        Hr = tbpm.Hamiltonian(s.num_orb, ip, ix, dt, precision=prec, handle=h)
        h.set_real_layout(False)
        Hc = tbpm.Hamiltonian(s.num_orb, ip, ix, dt, precision=prec, handle=h)
        assert Hr.info()["real_values"] and not Hc.info()["real_values"]
        a = tbpm.corr_dos(Hr, bes, psi, 20, chk_thr=1e-4)
        b = tbpm.corr_dos(Hc, bes, psi, 20, chk_thr=1e-4)
        np.testing.assert_array_equal(a, b)
        Hr.close()
        Hc.close()
    finally:
        h.close()


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_tile_order_equals_caller_order(gpu, prec):
    """Each row keeps its entries' order, so the states are the same; only
    the FP64 correlation sums run over the rows in another order."""
    s = tb.make_graphene_sample(40, vacancies=tb.random_vacancies(2 * 40 * 40, 0.02, 4))
    tb.add_anderson_disorder(s, 0.7, 5)
    s.rescale_ham(9.0)
    ip, ix, dt = _csr(s)
    bes = tbpm.bessel_series(tau_of(s), 1e-15)
    psi = _states(s.num_orb, 4)
    a = tbpm.corr_dos(tbpm.Hamiltonian(s.num_orb, ip, ix, dt, precision=prec), bes, psi, 40)
    for order in (s.tile_order(1.0), _random_order(s.num_orb, 3)):
        H = tbpm.Hamiltonian(s.num_orb, ip, ix, dt, precision=prec, order=order)
        b = tbpm.corr_dos(H, bes, psi, 40)
        assert np.max(np.abs(a - b)) < 1e-13
        H.close()


def _same_correlations(a, b, exact):
    """Blocked and one-term correlations: identical when the step ends with
    a single term (same partial sums); when it ends with a pass of two or
    more terms the FP64 sums are grouped by tile instead of by chunk, a few
    ulps."""
    if exact:
        np.testing.assert_array_equal(a, b)
    else:
        assert np.max(np.abs(a - b)) < 1e-13


def _blocked_sample(L=48):
    s = tb.make_graphene_sample(L, vacancies=tb.random_vacancies(2 * L * L, 0.02, 11))
    tb.add_anderson_disorder(s, 0.6, 12)
    s.rescale_ham(9.0)
    return s


@pytest.mark.parametrize("prec,K", [("fp32", 8), ("fp32", 3), ("fp32", 1), ("fp64", 1),
                                    ("fp64", 4)])
@pytest.mark.parametrize("nb_prec", [1e-15, 1e-14, 1e-6])
@pytest.mark.parametrize("terms", [3, 4, 5])
def test_three_term_pass_bitwise(gpu, prec, K, nb_prec, terms):
    """Three, four or five terms per pass over spatial tiles give the same states,
    bit for bit, as one term per pass, and the same correlations up to the
    grouping of their final FP64 sums; series lengths 21, 20 and 12 end the
    step with a pair pass, one single term and a pair pass (three terms per
    pass), or a full pass, a triple and a triple (four)."""
    s = _blocked_sample()
    ip, ix, dt = _csr(s)
    order, tiles = s.tile_order(0.8, return_tiles=True)
    bes = tbpm.bessel_series(tau_of(s), nb_prec)
    psi = _states(s.num_orb, K, seed=21)
    h = tb._lib.Handle(0)
    try:
        h.set_blocking(True, terms=terms)
        H = tbpm.Hamiltonian(s.num_orb, ip, ix, dt, precision=prec, handle=h, order=order,
                             tiles=tiles)
        nblk, halo = H.blocking_info(K)
        assert nblk > 1 and 1.0 < halo < 5.0
        a = tbpm.corr_dos(H, bes, psi, 30, chk_thr=1e-4)
        sa = tbpm.cheb_step(H, bes, psi, True)
        h.set_blocking(False)
        assert H.blocking_info(K) == (0, 0.0)
        b = tbpm.corr_dos(H, bes, psi, 30, chk_thr=1e-4)
        np.testing.assert_array_equal(sa, tbpm.cheb_step(H, bes, psi, True))
        _same_correlations(a, b, exact=(len(bes) - terms - 1) % terms == 1)
        ref = oracle.orc_tbpm_dos(ip, ix, dt, bes, psi, 30)
        assert np.max(np.abs(a - ref)) < (1e-12 if prec == "fp64" else 2e-5)
        H.close()
    finally:
        h.close()


@pytest.mark.parametrize("prec,K", [("fp64", 4), ("fp32", 6)])
@pytest.mark.parametrize("terms", [3, 4, 5])
def test_three_term_pass_scattered_tiles(gpu, prec, K, terms):
    """Tiles that are not spatial (a random row order cut every 16 rows):
    the halo rings then reach far across the sample, which the tables must
    still get right; the states stay bit-identical to one term per pass
    (the correlations too where the step ends with a single term)."""
    s = tb.make_graphene_sample(12, vacancies=tb.random_vacancies(2 * 12 * 12, 0.05, 6))
    tb.add_anderson_disorder(s, 0.9, 7)
    s.set_magnetic_field(25.0, 0)
    s.rescale_ham(9.0)
    ip, ix, dt = _csr(s)
    n = s.num_orb
    bes = tbpm.bessel_series(tau_of(s), 1e-14)
    psi = _states(n, K, seed=23)
    h = tb._lib.Handle(0)
    try:
        h.set_blocking(True, 16, terms=terms)
        H = tbpm.Hamiltonian(n, ip, ix, dt, precision=prec, handle=h,
                             order=_random_order(n, 12), tiles=_row_tiles(n, 16))
        nblk, halo = H.blocking_info(K)
        assert nblk == (n + 15) // 16 and halo > 2.0
        a = tbpm.corr_dos(H, bes, psi, 25, chk_thr=1e-4)
        sa = tbpm.cheb_step(H, bes, psi, True)
        h.set_blocking(False)
        b = tbpm.corr_dos(H, bes, psi, 25, chk_thr=1e-4)
        np.testing.assert_array_equal(sa, tbpm.cheb_step(H, bes, psi, True))
        _same_correlations(a, b, exact=(len(bes) - terms - 1) % terms == 1)
        ref = oracle.orc_tbpm_dos(ip, ix, dt, bes, psi, 25)
        assert np.max(np.abs(a - ref)) < (1e-12 if prec == "fp64" else 2e-5)
        H.close()
    finally:
        h.close()


@pytest.mark.parametrize("terms", [3, 4, 5])
def test_three_term_pass_wide_rows(gpu, terms):
    """Rows of ~9 entries (a random sparse Hermitian matrix, wider than the
    lattice rows the pipelined phase is specialised for)."""
    rng = np.random.default_rng(31)
    n = 400
    m = 4 * n
    hi = rng.integers(0, n, m)
    hj = rng.integers(0, n, m)
    keep = hi != hj
    hi, hj = hi[keep], hj[keep]
    hv = (rng.normal(size=len(hi)) + 1j * rng.normal(size=len(hi))) * 0.05
    orb = rng.uniform(-0.2, 0.2, n)
    ip, ix, dt = oracle.build_ham_csr(orb, hi, hj, hv)
    assert np.max(np.diff(ip)) > 6
    bes = tbpm.bessel_series(math.pi, 1e-14)
    psi = _states(n, 2, seed=24)
    h = tb._lib.Handle(0)
    try:
        h.set_blocking(True, 64, terms=terms)
        H = tbpm.Hamiltonian(n, ip, ix, dt, precision="fp64", handle=h,
                             tiles=_row_tiles(n, 32))
        assert H.blocking_info(2)[0] > 1
        a = tbpm.corr_dos(H, bes, psi, 20)
        sa = tbpm.cheb_step(H, bes, psi, True)
        h.set_blocking(False)
        b = tbpm.corr_dos(H, bes, psi, 20)
        np.testing.assert_array_equal(sa, tbpm.cheb_step(H, bes, psi, True))
        _same_correlations(a, b, exact=(len(bes) - terms - 1) % terms == 1)
        ref = oracle.orc_tbpm_dos(ip, ix, dt, bes, psi, 20)
        assert np.max(np.abs(a - ref)) < 1e-12
        H.close()
    finally:
        h.close()


@pytest.mark.parametrize("prec,K,field", [("fp32", 8, 0.0), ("fp32", 16, 0.0), ("fp64", 8, 0.0),
                                          ("fp64", 2, 0.0), ("fp64", 8, 30.0)])
@pytest.mark.parametrize("terms", [3, 4, 5])
def test_three_term_pass_workgroup_sizes(gpu, prec, K, field, terms):
    """Tiles of the benchmark's size (3.5 nm on 131,072-orbital graphene;
    3.0 nm with four terms per pass, whose three halo rings would not fit
    the LDS at 128 bytes per row):
    the tile-pass kernel runs with 16 waves per workgroup where the widest
    tile gives every lane (K/2 per row) two rows, else with 8, and always
    with 8 for complex values in complex128 (registers); the states stay
    bit-identical to one term per pass either way (the step of 21 terms
    ends with a pair pass, or a full pass of four terms)."""
    s = tb.make_graphene_sample(256)
    tb.add_anderson_disorder(s, 0.4, 15)
    if field:
        s.set_magnetic_field(field, 0)
    s.rescale_ham(9.0)
    ip, ix, dt = _csr(s)
    n = s.num_orb
    order, tiles = s.tile_order(3.5 if terms == 3 else 3.0, return_tiles=True)
    bes = tbpm.bessel_series(tau_of(s), 1e-15)
    psi = _states(n, K, seed=27)
    h = tb._lib.Handle(0)
    try:
        h.set_blocking(True, terms=terms)
        H = tbpm.Hamiltonian(n, ip, ix, dt, precision=prec, handle=h, order=order,
                             tiles=tiles)
        assert H.info()["real_values"] == (field == 0.0)
        rows, nt = H.blocking_launch(K)
        wide = prec == "fp32" or field == 0.0
        assert rows > 0 and nt == (1024 if wide and rows * (K // 2) >= 2048 else 512)
        if (prec, K) == ("fp32", 16) or (prec, K, field) == ("fp64", 2, 0.0):
            assert nt == (1024 if K == 16 else 512)     # both sizes are exercised
        a = tbpm.corr_dos(H, bes, psi, 12)
        sa = tbpm.cheb_step(H, bes, psi, True)
        h.set_blocking(False)
        assert H.blocking_launch(K) == (0, 0)
        b = tbpm.corr_dos(H, bes, psi, 12)
        np.testing.assert_array_equal(sa, tbpm.cheb_step(H, bes, psi, True))
        _same_correlations(a, b, exact=False)
        ref = oracle.orc_tbpm_dos(ip, ix, dt, bes, psi[:1], 12)
        assert np.max(np.abs(a[:1] - ref)) < (1e-12 if prec == "fp64" else 2e-5)
        H.close()
    finally:
        h.close()


@pytest.mark.parametrize("terms", [3, 4, 5])
def test_three_term_pass_graph_and_step(gpu, terms):
    """hipGraph replay of blocked steps equals launched steps; one blocked
    step (cheb_step) equals the oracle's step."""
    s = _blocked_sample(32)
    ip, ix, dt = _csr(s)
    order, tiles = s.tile_order(0.7, return_tiles=True)
    bes = tbpm.bessel_series(tau_of(s), 1e-15)
    psi = _states(s.num_orb, 4, seed=22)
    h = tb._lib.Handle(0)
    try:
        h.set_blocking(True, terms=terms)
        H = tbpm.Hamiltonian(s.num_orb, ip, ix, dt, precision="fp32", handle=h, order=order,
                             tiles=tiles)
        assert H.blocking_info(4)[0] > 1
        h.set_graph(True)
        a = tbpm.corr_dos(H, bes, psi, 70)
        h.set_graph(False)
        b = tbpm.corr_dos(H, bes, psi, 70)
        np.testing.assert_array_equal(a, b)
        H64 = tbpm.Hamiltonian(s.num_orb, ip, ix, dt, precision="fp64", handle=h,
                               order=order, tiles=tiles)
        assert H64.blocking_info(3)[0] > 1
        g = tbpm.cheb_step(H64, bes, psi[:3], True)
        for k in range(3):
            r = oracle.orc_step(ip, ix, dt, bes, psi[k], True)
            assert np.max(np.abs(g[k] - r)) < 1e-14
        H.close()
        H64.close()
    finally:
        h.close()


@pytest.mark.parametrize("terms", [3, 4, 5])
def test_block_tiles_split_to_fit(gpu, terms):
    """Tiles too large for the LDS at this row size (8 nm tiles, 16
    complex128 states: 256 bytes a row) are split until each block-tile's
    rings fit; the states stay bit-identical to one term per pass."""
    s = _blocked_sample(64)
    ip, ix, dt = _csr(s)
    order, tiles = s.tile_order(8.0, return_tiles=True)
    bes = tbpm.bessel_series(tau_of(s), 1e-15)
    psi = _states(s.num_orb, 16, seed=41)
    h = tb._lib.Handle(0)
    try:
        h.set_blocking(True, terms=terms)
        H = tbpm.Hamiltonian(s.num_orb, ip, ix, dt, precision="fp64", handle=h, order=order,
                             tiles=tiles)
        nblk, halo = H.blocking_info(16)
        assert nblk > len(tiles) - 1 and halo > 1.0
        rows, nt = H.blocking_launch(16)
        assert 0 < rows and rows * 256 <= 160 * 1024
        a = tbpm.corr_dos(H, bes, psi, 10)
        sa = tbpm.cheb_step(H, bes, psi, True)
        h.set_blocking(False)
        b = tbpm.corr_dos(H, bes, psi, 10)
        np.testing.assert_array_equal(sa, tbpm.cheb_step(H, bes, psi, True))
        _same_correlations(a, b, exact=(len(bes) - terms - 1) % terms == 1)
        ref = oracle.orc_tbpm_dos(ip, ix, dt, bes, psi[:2], 10)
        assert np.max(np.abs(a[:2] - ref)) < 1e-12
        H.close()
    finally:
        h.close()


@pytest.mark.parametrize("prec", ["fp32", "fp64"])
@pytest.mark.parametrize("terms", [4, 5])
def test_four_term_pass_long_tiles(gpu, prec, terms):
    """12 nm tiles with one state (8 or 16 bytes a row): a lane owns more
    rows than the four output partials it holds in registers between the
    third and fourth phase of a pass (complex64: ~5,500 owned rows for
    1024 lanes), so the rest go through memory; the states stay
    bit-identical to one term per pass."""
    s = _blocked_sample(128)
    ip, ix, dt = _csr(s)
    order, tiles = s.tile_order(12.0, return_tiles=True)
    bes = tbpm.bessel_series(tau_of(s), 1e-15)
    psi = _states(s.num_orb, 1, seed=43)
    h = tb._lib.Handle(0)
    try:
        h.set_blocking(True, terms=terms)
        H = tbpm.Hamiltonian(s.num_orb, ip, ix, dt, precision=prec, handle=h, order=order,
                             tiles=tiles)
        rows, nt = H.blocking_launch(1)
        assert nt == 1024 and rows > (4500 if prec == "fp32" else 2000)
        a = tbpm.corr_dos(H, bes, psi, 10)
        sa = tbpm.cheb_step(H, bes, psi, True)
        h.set_blocking(False)
        b = tbpm.corr_dos(H, bes, psi, 10)
        np.testing.assert_array_equal(sa, tbpm.cheb_step(H, bes, psi, True))
        _same_correlations(a, b, exact=False)
        ref = oracle.orc_tbpm_dos(ip, ix, dt, bes, psi, 10)
        assert np.max(np.abs(a - ref)) < (1e-12 if prec == "fp64" else 2e-5)
        H.close()
    finally:
        h.close()


def test_order_must_be_a_permutation(gpu):
    s = graphene_sample(4)
    ip, ix, dt = _csr(s)
    bad = np.arange(s.num_orb, dtype=np.int64)
    bad[3] = 2
    with pytest.raises(tb._lib.TBGError, match="permutation"):
        tbpm.Hamiltonian(s.num_orb, ip, ix, dt, order=bad)
    bad[3] = s.num_orb
    with pytest.raises(tb._lib.TBGError, match="permutation"):
        tbpm.Hamiltonian(s.num_orb, ip, ix, dt, order=bad)


def test_batch_ordered_and_plain(gpu):
    hams, psis, refs = [], [], []
    bes = None
    for i, L in enumerate((9, 14, 6)):
        s = tb.make_graphene_sample(L)
        tb.add_anderson_disorder(s, 0.4, 30 + i)
        s.rescale_ham(9.0)
        ip, ix, dt = _csr(s)
        bes = tbpm.bessel_series(tau_of(s), 1e-15)
        H = tbpm.Hamiltonian(s.num_orb, ip, ix, dt, precision="fp64",
                             order=s.tile_order(1.0) if i != 1 else None)
        psi = _states(s.num_orb, 2, seed=300 + i)
        hams.append(H)
        psis.append(psi)
        refs.append(oracle.orc_tbpm_dos(ip, ix, dt, bes, psi, 24))
    out = tbpm.corr_dos_batch(hams, bes, psis, 24)
    for o, r in zip(out, refs):
        assert np.max(np.abs(o - r)) < 1e-12


def test_batch_mixed_real_and_complex(gpu):
    """A batch of a real and a complex (Peierls) Hamiltonian widens the real
    values to complex; each instance still matches the oracle."""
    hams, psis, refs = [], [], []
    bes = None
    for i, L in enumerate((10, 12, 7)):
        s = tb.make_graphene_sample(L)
        if i == 1:
            s.set_magnetic_field(30.0, 0)
        s.rescale_ham(9.0)
        ip, ix, dt = _csr(s)
        bes = tbpm.bessel_series(tau_of(s), 1e-15)
        H = tbpm.Hamiltonian(s.num_orb, ip, ix, dt, precision="fp64")
        assert H.info()["real_values"] == (i != 1)
        psi = _states(s.num_orb, 2, seed=200 + i)
        hams.append(H)
        psis.append(psi)
        refs.append(oracle.orc_tbpm_dos(ip, ix, dt, bes, psi, 24))
    out = tbpm.corr_dos_batch(hams, bes, psis, 24)
    for o, r in zip(out, refs):
        assert np.max(np.abs(o - r)) < 1e-12


def test_dos_run_pieces_equal_one_call(gpu):
    s = graphene_sample(12)
    ip, ix, dt = _csr(s)
    bes = tbpm.bessel_series(tau_of(s), 1e-15)
    psi = _states(s.num_orb, 2)
    H = tbpm.Hamiltonian(s.num_orb, ip, ix, dt, precision="fp64")
    full = tbpm.corr_dos(H, bes, psi, 50)
    run = tbpm.DosRun(H, bes, psi, 50)
    ms = run.advance(17) + run.advance(33)
    assert ms > 0
    np.testing.assert_array_equal(run.result(), full)
    run.close()


@pytest.mark.parametrize("ordered", [False, True])
def test_cheb_step_and_poly(gpu, ordered):
    s = tb.make_graphene_sample(10)
    tb.add_anderson_disorder(s, 0.5, 9)
    s.set_magnetic_field(15.0, 1)
    s.rescale_ham(9.0)
    ip, ix, dt = _csr(s)
    bes = tbpm.bessel_series(tau_of(s), 1e-15)
    psi = _states(s.num_orb, 3)
    H = tbpm.Hamiltonian(s.num_orb, ip, ix, dt, precision="fp64",
                         order=_random_order(s.num_orb, 9) if ordered else None)
    for fwd in (True, False):
        g = tbpm.cheb_step(H, bes, psi, fwd)
        for k in range(3):
            r = oracle.orc_step(ip, ix, dt, bes, psi[k], fwd)
            assert np.max(np.abs(g[k] - r)) < 1e-14
    coef = golden("fermi.npz")["coef_F"][:400]
    g = tbpm.cheb_poly(H, coef, psi)
    # dense evaluation of sum_k c_k T_k(H) psi
    n = s.num_orb
    Hd = np.zeros((n, n), complex)
    for r in range(n):
        Hd[r, ix[ip[r]:ip[r + 1]]] += dt[ip[r]:ip[r + 1]]
    w, v = np.linalg.eigh(Hd)
    f = np.polynomial.chebyshev.chebval(w, coef)
    dense = (v * f) @ (v.conj().T @ psi.T)
    assert np.max(np.abs(g - dense.T)) < 1e-12


def test_cfg1_full_size_fp64_vs_oracle_and_kspace(gpu):
    """The reference CPU case at full size: 256x256 graphene (131,072
    orbitals), 1024 steps, Bessel terms to 1e-15."""
    L = 256
    s = graphene_sample(L)
    ip, ix, dt = [t.cpu().numpy() for t in s.build_ham_csr_device()]
    bes = tbpm.bessel_series(tau_of(s), 1e-15)
    assert len(bes) == 21
    psi = _states(s.num_orb, 1)
    T = 1024
    H = tbpm.Hamiltonian(s.num_orb, ip, ix, dt, precision="fp64")
    c = tbpm.corr_dos(H, bes, psi, T)[0]
    ref = oracle.orc_tbpm_dos(ip, ix, dt, bes, psi, T)[0]
    assert np.max(np.abs(c - ref)) < 1e-10
    pts = [0, 1, 2, 10, 100, 511, 1024]
    k = oracle.kspace_graphene_corr(L, L, s.hop_v[0], psi[0], tau_of(s), pts)
    assert np.max(np.abs(c[pts] - k)) < 1e-10
    H32 = tbpm.Hamiltonian(s.num_orb, ip, ix, dt, precision="fp32")
    c32 = tbpm.corr_dos(H32, bes, psi, T, chk_thr=1e-4)[0]
    assert np.max(np.abs(c32 - ref)) < 2e-4


def test_large_kspace_spot_check(gpu):
    """2048x2048 (8.4M orbitals, the benchmark sample), 8 states, 64 steps,
    FP64 and FP32 against exact k-space propagation at sampled steps."""
    L = 2048
    s = graphene_sample(L)
    ip, ix, dt = s.build_ham_csr_device()
    bes = tbpm.bessel_series(tau_of(s), 1e-15)
    psi = _states(s.num_orb, 8)
    T = 64
    pts = [0, 1, 7, 33, 64]
    ex = np.array([oracle.kspace_graphene_corr(L, L, s.hop_v[0], psi[k], tau_of(s), pts)
                   for k in range(8)])
    for prec, tol in (("fp64", 1e-11), ("fp32", 2e-5)):
        H = tbpm.Hamiltonian(s.num_orb, ip, ix, dt, precision=prec)
        c = tbpm.corr_dos(H, bes, psi, T, chk_thr=1e-4)
        H.close()
        assert np.max(np.abs(c[:, pts] - ex)) < tol, prec


def test_norm_check_raises(gpu):
    """rescale below the spectral radius makes the expansion diverge; the
    reference STOPs (funcs.f90:175-192), here NormError is raised."""
    s = graphene_sample(8, rescale=2.0)
    ip, ix, dt = _csr(s)
    bes = tbpm.bessel_series(tau_of(s), 1e-14)
    psi = _states(s.num_orb, 1)
    H = tbpm.Hamiltonian(s.num_orb, ip, ix, dt, precision="fp64")
    from tbplas_amd._lib import NormError
    with pytest.raises(NormError) as ei:
        tbpm.corr_dos(H, bes, psi, 40, chk_step=16)
    assert ei.value.step == 16


def test_batch_equals_individual(gpu):
    rng = np.random.default_rng(0)
    hams, psis, refs = [], [], []
    bes = None
    for i, L in enumerate((8, 13, 5, 16)):
        n0 = 2 * L * L
        s = tb.make_graphene_sample(L, vacancies=tb.random_vacancies(n0, 0.03 * (i % 2), i))
        tb.add_anderson_disorder(s, 0.5 * i, 10 + i)
        s.rescale_ham(9.0)
        ip, ix, dt = _csr(s)
        bes = tbpm.bessel_series(tau_of(s), 1e-15)
        H = tbpm.Hamiltonian(s.num_orb, ip, ix, dt, precision="fp64")
        psi = _states(s.num_orb, 2, seed=100 + i)
        hams.append(H)
        psis.append(psi)
        refs.append(oracle.orc_tbpm_dos(ip, ix, dt, bes, psi, 24))
    out = tbpm.corr_dos_batch(hams, bes, psis, 24)
    for o, r in zip(out, refs):
        assert np.max(np.abs(o - r)) < 1e-12


@pytest.mark.parametrize("prec,K,terms", [("fp64", 2, 3), ("fp32", 3, 3), ("fp64", 1, 3),
                                          ("fp64", 2, 4), ("fp32", 8, 4)])
def test_batch_three_term_pass(gpu, prec, K, terms):
    """A batch whose instances all have spatial tiles runs three terms per
    pass over the concatenation (workgroups of at most 48 rows, never across
    two instances); every instance matches the oracle and the one-term-per-
    pass batch, and a second run reuses the cached tables."""
    hams = {True: [], False: []}
    psis, refs = [], []
    bes = None
    handles = {}
    for blocked in (True, False):
        hd = tb._lib.Handle(0)
        hd.set_blocking(blocked, 48, terms=terms)
        handles[blocked] = hd
    for i, L in enumerate((9, 14, 5, 11)):
        n0 = 2 * L * L
        s = tb.make_graphene_sample(L, vacancies=tb.random_vacancies(n0, 0.02 * (i % 2), 40 + i))
        tb.add_anderson_disorder(s, 0.3 * i, 50 + i)
        if i == 3:
            s.set_magnetic_field(25.0, 0)
        s.rescale_ham(9.0)
        ip, ix, dt = _csr(s)
        bes = tbpm.bessel_series(tau_of(s), 1e-15)
        order, tiles = s.tile_order(0.6, return_tiles=True)
        for blocked in (True, False):
            hams[blocked].append(tbpm.Hamiltonian(s.num_orb, ip, ix, dt, precision=prec,
                                                  handle=handles[blocked], order=order,
                                                  tiles=tiles))
        psi = _states(s.num_orb, K, seed=400 + i)
        psis.append(psi)
        refs.append(oracle.orc_tbpm_dos(ip, ix, dt, bes, psi, 20))
    tol = 1e-12 if prec == "fp64" else 2e-5
    out_b = tbpm.corr_dos_batch(hams[True], bes, psis, 20)
    out_u = tbpm.corr_dos_batch(hams[False], bes, psis, 20)
    again = tbpm.corr_dos_batch(hams[True], bes, psis, 20)
    for ob, ou, oa, r in zip(out_b, out_u, again, refs):
        assert ob.shape == r.shape
        assert np.max(np.abs(ob - r)) < tol
        assert np.max(np.abs(ob - ou)) < (1e-13 if prec == "fp64" else 2e-6)
        assert np.array_equal(ob, oa)
    for blocked in (True, False):
        for H in hams[blocked]:
            H.close()
        handles[blocked].close()


def _random_order(n, seed):
    return np.random.default_rng(seed).permutation(n).astype(np.int64)


def _row_tiles(n, rows):
    return np.r_[np.arange(0, n, rows), n].astype(np.int64)


@pytest.mark.parametrize("layout", ["plain", "ordered", "tiles", "tiles4", "tiles5"])
def test_ac_reference_fixture(gpu, layout):
    """ordered: the device layout stores the rows in a random order (every
    gather leaves its slice); tiles: three terms per pass (tiles4: four,
    tiles5: five), workgroups of 32 rows; results stay in the caller's
    order."""
    g = golden("ref_ac.npz")
    n = len(g["indptr"]) - 1
    tiled = layout.startswith("tiles")
    hd = tb._lib.Handle(0) if tiled else None
    if hd is not None:
        hd.set_blocking(True, 32, terms={"tiles4": 4, "tiles5": 5}.get(layout, 3))
    H = tbpm.Hamiltonian(n, g["indptr"], g["indices"], g["hop"], g["dx"], g["dy"],
                         cur_scale=float(g["rescale"]), precision="fp64", handle=hd,
                         order=_random_order(n, 7) if layout == "ordered" else None,
                         tiles=_row_tiles(n, 32) if tiled else None)
    if tiled:
        assert H.blocking_info(3 * g["psi0"].shape[0])[0] == 4
    cF = tbpm.fermi_cheb_coef(2 ** 15, float(g["beta_re"]), float(g["mu_re"]), False, 1e-12)
    cO = tbpm.fermi_cheb_coef(2 ** 15, float(g["beta_re"]), float(g["mu_re"]), True, 1e-12)
    c = tbpm.corr_ac(H, g["bes"], cF, cO, g["psi0"], int(g["T"]))
    S = c.shape[0]
    avg = np.zeros_like(c[0])
    for s in range(S):
        avg = avg + c[s] / S
    ref = g["corr"]
    assert np.all(avg[:, 0] == 0)
    assert np.max(np.abs(avg - ref)) < 1e-11 * max(1.0, np.max(np.abs(ref)))
    H.close()
    if hd is not None:
        hd.close()


def test_ac_fp32_vs_oracle(gpu):
    s = tb.make_graphene_sample(10)
    s.set_magnetic_field(10.0, 0)
    s.rescale_ham(9.0)
    ip, ix, h, dx, dy = oracle.build_ham_dxy_loop(s.orb_eng, s.hop_i, s.hop_j, s.hop_v, s.dr)
    bes = tbpm.bessel_series(tau_of(s, 0.5), 1e-14)
    beta_re = 1 / (8.617333262e-5 * 300) * 9.0
    cF = tbpm.fermi_cheb_coef(2 ** 15, beta_re, 0.02, False, 1e-12)
    cO = tbpm.fermi_cheb_coef(2 ** 15, beta_re, 0.02, True, 1e-12)
    psi = _states(s.num_orb, 3)
    T = 12
    ref = oracle.orc_tbpm_accond(ip, ix, h, dx, dy, 9.0, beta_re, 0.02, bes, psi, T)
    scale = np.max(np.abs(ref))
    for prec, tol in (("fp64", 1e-11), ("fp32", 1e-4)):
        H = tbpm.Hamiltonian(s.num_orb, ip, ix, h, dx, dy, cur_scale=9.0, precision=prec)
        c = tbpm.corr_ac(H, bes, cF, cO, psi, T, chk_thr=1e-4)
        assert np.max(np.abs(c - ref)) < tol * scale, prec


def test_fermi_coef_product_vs_reference_fixture(gpu):
    g = golden("fermi.npz")
    a = tbpm.fermi_cheb_coef(int(g["nr_fermi"]), float(g["beta_re"]), 0.0, False, 1e-12)
    b = tbpm.fermi_cheb_coef(int(g["nr_fermi"]), float(g["beta_re"]), 0.0, True, 1e-12)
    assert len(a) == len(g["coef_F"]) and len(b) == len(g["coef_omF"])
    assert np.max(np.abs(a - g["coef_F"])) < 1e-14
    assert np.max(np.abs(b - g["coef_omF"])) < 1e-14
    c = tbpm.fermi_cheb_coef(int(g["nr_fermi"]), float(g["beta_re_cold"]),
                             float(g["mu_re_cold"]), False, 1e-12)
    assert len(c) == len(g["coef_F_cold"])
    assert np.max(np.abs(c - g["coef_F_cold"])) < 1e-14


@pytest.mark.parametrize("layout", ["plain", "ordered", "tiles", "tiles4", "tiles5"])
def test_dc_reference_fixture(gpu, layout):
    g = golden("ref_dc.npz")
    n = len(g["indptr"]) - 1
    tiled = layout.startswith("tiles")
    hd = tb._lib.Handle(0) if tiled else None
    if hd is not None:
        hd.set_blocking(True, 32, terms={"tiles4": 4, "tiles5": 5}.get(layout, 3))
    H = tbpm.Hamiltonian(n, g["indptr"], g["indices"], g["hop"], g["dx"], g["dy"],
                         cur_scale=float(g["rescale"]), precision="fp64", handle=hd,
                         order=_random_order(n, 8) if layout == "ordered" else None,
                         tiles=_row_tiles(n, 32) if tiled else None)
    if tiled:
        assert H.blocking_info(2 * g["psi0"].shape[0])[0] == 4
    dos, dc = tbpm.corr_dc(H, g["bes"], g["psi0"], int(g["T"]), float(g["t_step"]),
                           g["energies"], g["en_inds"])
    S = dos.shape[0]
    ne = len(g["en_inds"])
    adc = np.zeros_like(dos[0])
    acc = np.zeros_like(g["dc"])
    for s in range(S):
        adc = adc + dos[s] / S
        acc[:, :ne, :] = acc[:, :ne, :] + dc[s] / S
    assert np.max(np.abs(adc - g["dos"])) < 1e-12
    assert np.max(np.abs(acc - g["dc"])) < 1e-10 * max(1.0, np.max(np.abs(g["dc"])))
    H.close()
    if hd is not None:
        hd.close()


def test_dc_fp32_vs_oracle(gpu):
    s = tb.make_graphene_sample(10, vacancies=tb.random_vacancies(200, 0.02, 8))
    tb.add_anderson_disorder(s, 0.6, 8)
    s.rescale_ham(9.0)
    ip, ix, h, dx, dy = oracle.build_ham_dxy_loop(s.orb_eng, s.hop_i, s.hop_j, s.hop_v, s.dr)
    tau = tau_of(s)
    bes = tbpm.bessel_series(tau, 1e-14)
    T = 12
    en = np.array([0.5 * i * 18.0 / T - 9.0 for i in range(2 * T)])
    inds = np.array([5, 11, 12, 13, 20])
    psi = _states(s.num_orb, 4)
    t_step = 2 * math.pi / 18.0
    rd, rc = oracle.orc_tbpm_dccond(ip, ix, h, dx, dy, 9.0, bes, psi, T, t_step, en, inds)
    for prec, tol in (("fp64", 1e-10), ("fp32", 2e-4)):
        H = tbpm.Hamiltonian(s.num_orb, ip, ix, h, dx, dy, cur_scale=9.0, precision=prec)
        d, c = tbpm.corr_dc(H, bes, psi, T, t_step, en, inds, chk_thr=1e-4)
        assert np.max(np.abs(d - rd)) < tol, prec
        assert np.max(np.abs(c - rc)) < tol * max(1.0, np.max(np.abs(rc))), prec


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-13), ("fp32", 1e-5)])
def test_diagonal_hamiltonian_closed_form(gpu, prec, tol):
    """H = diag(e) (no hoppings): C(t) = sum_i |psi_i|^2 exp(-i e_i tau t);
    n = 1 and n = 5, with and without three terms per pass."""
    tau = 1.3
    bes = tbpm.bessel_series(tau, 1e-15)
    for e in (np.array([0.37]), np.array([-0.9, -0.2, 0.0, 0.45, 0.8])):
        n = len(e)
        ip = np.arange(n + 1, dtype=np.int64)
        ix = np.arange(n, dtype=np.int64)
        rng = np.random.default_rng(n)
        psi = rng.normal(size=(2, n)) + 1j * rng.normal(size=(2, n))
        psi /= np.linalg.norm(psi, axis=1, keepdims=True)
        T = 40
        t = np.arange(T + 1)
        exact = (np.abs(psi) ** 2) @ np.exp(-1j * np.outer(e, t) * tau)
        for tiles in (None, np.array([0, n], dtype=np.int64)):
            H = tbpm.Hamiltonian(n, ip, ix, e.astype(np.complex128), precision=prec,
                                 tiles=tiles)
            c = tbpm.corr_dos(H, bes, psi, T)
            assert np.max(np.abs(c - exact)) < tol
            H.close()


def test_state_count_limits_and_one_step(gpu):
    s = graphene_sample(8)
    ip, ix, dt = _csr(s)
    bes = tbpm.bessel_series(tau_of(s), 1e-15)
    H = tbpm.Hamiltonian(s.num_orb, ip, ix, dt, precision="fp32")
    psi = _states(s.num_orb, 64, seed=40)
    many = tbpm.corr_dos(H, bes, psi, 3)
    for k in (0, 31, 63):
        one = tbpm.corr_dos(H, bes, psi[k:k + 1], 3)[0]
        assert np.max(np.abs(one - many[k])) < 1e-6
    c1 = tbpm.corr_dos(H, bes, psi[:2], 1)
    assert c1.shape == (2, 2)
    assert np.max(np.abs(c1 - many[:2, :2])) < 1e-6
    with pytest.raises(tb._lib.TBGError, match=r"\[1, 64\]"):
        tbpm.corr_dos(H, bes, _states(s.num_orb, 65, seed=41), 3)
    with pytest.raises(tb._lib.TBGError, match="time steps"):
        tbpm.corr_dos(H, bes, psi[:1], 0)
    H.close()
