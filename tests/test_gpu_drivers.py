"""The remaining Chebyshev drivers on the same engine: LDOS correlation,
wave-function propagation psi(t) and quasi-eigenstates, against the oracle
(pinned in test_oracle.py against the reference's own subroutines and
closed forms) on identical states, in both precisions, in the caller's row
order and with spatial tiles (three terms per pass); and the Solver
methods that drive them."""
import math

import numpy as np
import pytest

import oracle
import tbplas_amd as tb
from tbplas_amd import states, tbpm

pytestmark = pytest.mark.gpu
needs_ref = pytest.mark.skipif(not oracle.ref_available(),
                               reason="oracle/_ref not built (reference sources absent)")


def _sample(L=14, W=0.6, field=0.0, seed=3):
    s = tb.make_graphene_sample(L, vacancies=tb.random_vacancies(2 * L * L, 0.02, seed))
    tb.add_anderson_disorder(s, W, seed + 1)
    if field:
        s.set_magnetic_field(field, 0)
    s.rescale_ham(9.0)
    return s


def _ham(s, prec, layout, handle=None):
    ip, ix, dt = oracle.build_ham_csr(s.orb_eng, s.hop_i, s.hop_j, s.hop_v)
    kw = {}
    if layout.startswith("tiles"):
        kw["order"], kw["tiles"] = s.tile_order(0.8, return_tiles=True)
        if layout in ("tiles4", "tiles5") and handle is None:   # four or five terms per pass
            handle = tb._lib.Handle(0)
            handle.set_blocking(True, terms=int(layout[-1]))
    H = tbpm.Hamiltonian(s.num_orb, ip, ix, dt, precision=prec, handle=handle, **kw)
    if layout.startswith("tiles"):
        assert H.blocking_info(4)[0] > 1
    return H, (ip, ix, dt)


def _tol(prec):
    return 1e-12 if prec == "fp64" else 3e-5


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("layout", ["caller", "tiles", "tiles4", "tiles5"])
def test_ldos_vs_oracle(gpu, prec, layout):
    s = _sample(field=20.0 if layout == "caller" else 0.0)
    H, (ip, ix, dt) = _ham(s, prec, layout)
    bes = tbpm.bessel_series(2 * math.pi / s.energy_range * s.rescale, 1e-15)
    psi = tbpm.ldos_states(states.random_states(s.num_orb, states.sample_seeds(11, 4)),
                           [0, 5, 17, 40, 41])
    assert np.allclose(np.linalg.norm(psi, axis=1), 1.0)
    T = 40
    g = tbpm.corr_ldos(H, bes, psi, T, chk_thr=1e-4)
    r = oracle.orc_tbpm_ldos(ip, ix, dt, bes, psi, T)
    assert g.shape == (4, T + 1)
    assert np.max(np.abs(g - r)) < _tol(prec)
    H.close()


@needs_ref
def test_ldos_vs_reference_live(gpu):
    """Averaged over the reference's own random states: the reference's
    tbpm_ldos output."""
    s = _sample(L=8)
    H, (ip, ix, dt) = _ham(s, "fp64", "caller")
    bes = tbpm.bessel_series(2 * math.pi / s.energy_range * s.rescale, 1e-14)
    S, T, seed, sites = 3, 30, 7, [2, 9, 33]
    ref = oracle.ref_tbpm_ldos(sites, bes, ip, ix, dt, seed, T, S)
    psi = np.array([oracle.ref_random_state(s.num_orb, seed * (i + 1)) for i in range(S)])
    g = tbpm.corr_ldos(H, bes, tbpm.ldos_states(psi, sites), T)
    avg = np.zeros(T + 1, dtype=complex)
    for i in range(S):
        avg = avg + g[i] / S
    assert np.max(np.abs(avg - ref)) < 1e-12
    H.close()


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("layout", ["caller", "tiles", "tiles4", "tiles5"])
@pytest.mark.parametrize("tl", [[0, 1, 2, 25, 40], [40], [0], [3, 17, 39]])
def test_psi_t_vs_oracle(gpu, prec, layout, tl):
    s = _sample()
    H, (ip, ix, dt) = _ham(s, prec, layout)
    bes = tbpm.bessel_series(2 * math.pi / s.energy_range * s.rescale, 1e-15)
    p0 = states.random_states(s.num_orb, states.sample_seeds(5, 1))[0]
    T = 40
    g = tbpm.psi_t(H, bes, p0, tl, T, chk_step=20, chk_thr=1e-4)
    r = oracle.orc_psi_t(ip, ix, dt, bes, p0, tl, T)
    assert g.shape == (len(tl), s.num_orb)
    assert np.max(np.abs(g - r)) < _tol(prec)
    if 0 in tl:   # the initial state as stored (complex64 rounds it)
        stored = p0 if prec == "fp64" else p0.astype(np.complex64).astype(np.complex128)
        np.testing.assert_array_equal(g[tl.index(0)], stored)
    H.close()


def test_psi_t_rejects_bad_time_log(gpu):
    s = _sample(L=6)
    H, _ = _ham(s, "fp64", "caller")
    bes = tbpm.bessel_series(math.pi, 1e-14)
    p0 = states.random_states(s.num_orb, states.sample_seeds(5, 1))[0]
    for bad in ([3, 1], [0, 0], [-1], [11]):
        with pytest.raises(tb._lib.TBGError):
            tbpm.psi_t(H, bes, p0, bad, 10)
    H.close()


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("layout", ["caller", "tiles", "tiles4", "tiles5"])
def test_quasi_eigenstates_vs_oracle(gpu, prec, layout):
    s = _sample(L=10)
    H, (ip, ix, dt) = _ham(s, prec, layout)
    t_step = 2 * math.pi / s.energy_range
    bes = tbpm.bessel_series(t_step * s.rescale, 1e-14)
    psi = states.random_states(s.num_orb, states.sample_seeds(9, 3))
    T, en = 32, [-1.0, 0.0, 0.35, 2.0]
    g = tbpm.quasi_eigenstates(H, bes, psi, T, t_step, en, chk_thr=1e-4)
    r = oracle.orc_quasi_eigenstates(ip, ix, dt, bes, psi, T, t_step, en)
    assert g.shape == (3, len(en), s.num_orb)
    assert np.max(np.abs(g - r)) < _tol(prec) * max(1.0, np.max(r)) * 10
    assert np.allclose(g.sum(axis=2), 1.0, atol=1e-12 if prec == "fp64" else 1e-5)
    H.close()


def test_quasi_eigenstates_many_states_and_energies(gpu):
    """12 states x 9 energies = 108 quasi-eigenstates: their norms and
    overlaps are summed in tiles of up to 256 per workgroup, more sums
    than a workgroup has threads."""
    s = _sample(L=8)
    H, (ip, ix, dt) = _ham(s, "fp64", "caller")
    t_step = 2 * math.pi / s.energy_range
    bes = tbpm.bessel_series(t_step * s.rescale, 1e-14)
    psi = states.random_states(s.num_orb, states.sample_seeds(19, 12))
    T, en = 24, list(np.linspace(-2.0, 2.0, 9))
    g = tbpm.quasi_eigenstates(H, bes, psi, T, t_step, en)
    r = oracle.orc_quasi_eigenstates(ip, ix, dt, bes, psi, T, t_step, en)
    assert np.max(np.abs(g - r)) < 1e-11 * max(1.0, np.max(r))
    assert np.allclose(g.sum(axis=2), 1.0, atol=1e-12)
    H.close()


def _config(T, S, precision="fp64"):
    c = tb.Config()
    c.generic["nr_time_steps"] = T
    c.generic["nr_random_samples"] = S
    c.gpu["precision"] = precision
    c.gpu["autotune"] = False
    return c


def test_solver_ldos_qe_psi_t(gpu, tmp_path):
    s = _sample(L=10, W=0.0)
    T, S = 24, 3
    cfg = _config(T, S)
    cfg.LDOS["site_indices"] = [0, 1, 2, 50]
    cfg.quasi_eigenstates["energies"] = [-0.5, 0.0, 0.5]
    solver = tb.Solver(s, cfg, directory=str(tmp_path / "sim"), prefix="d", echo_details=False)
    ip, ix, dt = oracle.build_ham_csr(s.orb_eng, s.hop_i, s.hop_j, s.hop_v)
    t_step = 2 * math.pi / s.energy_range
    bes = oracle.bessel_series(t_step * s.rescale, cfg.generic["Bessel_precision"])
    psi = states.random_states(s.num_orb, states.sample_seeds(1337, S))

    ldos = solver.calc_corr_ldos()
    assert ldos.shape == (T + 1,) and (tmp_path / "sim" / "d.corr_LDOS.npy").exists()
    r = oracle.orc_tbpm_ldos(ip, ix, dt, bes, tbpm.ldos_states(psi, [0, 1, 2, 50]), T)
    ref = np.zeros(T + 1, dtype=complex)
    for i in range(S):
        ref = ref + r[i] / S
    assert np.max(np.abs(ldos - ref)) < 1e-12

    qe = solver.calc_quasi_eigenstates(save_data=True)
    assert qe.shape == (3, s.num_orb) and (tmp_path / "sim" / "d.qe.npy").exists()
    q = oracle.orc_quasi_eigenstates(ip, ix, dt, bes, psi, T, t_step, [-0.5, 0.0, 0.5])
    qref = np.zeros_like(q[0])
    for i in range(S):
        qref = qref + q[i] / S
    assert np.max(np.abs(qe - qref)) < 1e-12

    p0 = np.zeros(s.num_orb, dtype=complex)
    p0[7] = 2.0
    frames = solver.calc_psi_t(p0, [5, 0, 5, 24], save_data=True)
    assert frames.shape == (3, s.num_orb)
    pr = oracle.orc_psi_t(ip, ix, dt, bes, p0 / 2.0, [0, 5, 24], T)
    assert np.max(np.abs(frames - pr)) < 1e-12
    with pytest.raises(ValueError):
        solver.calc_psi_t(p0, [T + 1])


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("layout", ["caller", "tiles", "tiles4", "tiles5"])
def test_dyn_pol_vs_oracle(gpu, prec, layout):
    s = _sample(L=10, field=15.0 if layout == "caller" else 0.0)
    H, (ip, ix, dt) = _ham(s, prec, layout)
    tau = 0.5 * 2 * math.pi / s.energy_range * s.rescale
    bes = tbpm.bessel_series(tau, 1e-14)
    beta_re = 1.0 / (8.617333262e-5 * 300.0) * s.rescale
    cF = tbpm.fermi_cheb_coef(2 ** 15, beta_re, 0.0, False, 1e-12)
    cO = tbpm.fermi_cheb_coef(2 ** 15, beta_re, 0.0, True, 1e-12)
    psi = states.random_states(s.num_orb, states.sample_seeds(13, 3))
    qs = [[1.0, 0.5, 0.0], [0.0, 2.0, 0.0], [0.3, 0.0, 0.0]]
    T = 20
    g = tbpm.corr_dyn_pol(H, bes, cF, cO, psi, T, qs, s.orb_pos, chk_thr=1e-4)
    r = oracle.orc_tbpm_dyn_pol(ip, ix, dt, bes, cF, cO, psi, T, qs, s.orb_pos)
    assert g.shape == (3, 3, T)
    scale = max(1e-3, np.max(np.abs(r)))
    assert np.max(np.abs(g - r)) < (1e-12 if prec == "fp64" else 2e-5) * max(1.0, scale)
    H.close()


@needs_ref
def test_dyn_pol_vs_reference_live(gpu):
    s = _sample(L=6, W=0.0)
    ip, ix, dt, dx, dy = oracle.build_ham_dxy_loop(s.orb_eng, s.hop_i, s.hop_j, s.hop_v, s.dr)
    H = tbpm.Hamiltonian(s.num_orb, ip, ix, dt, precision="fp64")
    tau = 0.5 * 2 * math.pi / s.energy_range * s.rescale
    bes = tbpm.bessel_series(tau, 1e-14)
    beta_re, mu_re = 1.0 / (8.617333262e-5 * 300.0) * s.rescale, 0.0
    S, T, seed, qs = 2, 12, 5, [[0.9, 0.1, 0.0]]
    ref = oracle.ref_tbpm_dyn_pol(bes, beta_re, mu_re, ip, ix, dt, s.rescale, dx, dy, s.orb_pos,
                                  seed, T, S, qs)
    cF = tbpm.fermi_cheb_coef(2 ** 15, beta_re, mu_re, False, 1e-12)
    cO = tbpm.fermi_cheb_coef(2 ** 15, beta_re, mu_re, True, 1e-12)
    psi = np.array([oracle.ref_random_state(s.num_orb, seed * (i + 1)) for i in range(S)])
    g = tbpm.corr_dyn_pol(H, bes, cF, cO, psi, T, qs, s.orb_pos)
    avg = np.zeros_like(g[0])
    for i in range(S):
        avg = avg + g[i] / S
    assert np.max(np.abs(avg - ref)) < 1e-12 * max(1.0, np.max(np.abs(ref)))
    H.close()


def test_solver_dyn_pol(gpu, tmp_path):
    s = _sample(L=8, W=0.0)
    T, S = 12, 2
    cfg = _config(T, S)
    cfg.dyn_pol["q_points"] = [[1.0, 0.0, 0.0], [0.5, 0.5, 0.0]]
    solver = tb.Solver(s, cfg, directory=str(tmp_path / "sim"), prefix="p", echo_details=False)
    c = solver.calc_corr_dyn_pol()
    assert c.shape == (2, T) and c.dtype == np.float64
    assert (tmp_path / "sim" / "p.corr_dyn_pol.npy").exists()
    ip, ix, dt = oracle.build_ham_csr(s.orb_eng, s.hop_i, s.hop_j, s.hop_v)
    tau = 0.5 * 2 * math.pi / s.energy_range * s.rescale
    bes = oracle.bessel_series(tau, cfg.generic["Bessel_precision"])
    beta_re = cfg.generic["beta"] * s.rescale
    cF = oracle.orc_fermi_coef(2 ** 15, beta_re, 0.0, False, 1e-12)
    cO = oracle.orc_fermi_coef(2 ** 15, beta_re, 0.0, True, 1e-12)
    psi = states.random_states(s.num_orb, states.sample_seeds(1337, S))
    r = oracle.orc_tbpm_dyn_pol(ip, ix, dt, bes, cF, cO, psi, T, cfg.dyn_pol["q_points"],
                                s.orb_pos)
    ref = np.zeros_like(r[0])
    for i in range(S):
        ref = ref + r[i] / S
    assert np.max(np.abs(c - ref)) < 1e-12 * max(1.0, np.max(np.abs(ref)))


def _dxy_ham(s, prec="fp64"):
    ip, ix, dt, dx, dy = oracle.build_ham_dxy_loop(s.orb_eng, s.hop_i, s.hop_j, s.hop_v, s.dr)
    H = tbpm.Hamiltonian(s.num_orb, ip, ix, dt, dx, dy, cur_scale=s.rescale, precision=prec)
    return H, (ip, ix, dt, dx, dy)


@pytest.mark.parametrize("N,field,direction", [(16, 0.0, 1), (64, 25.0, 2), (300, 10.0, 1)])
def test_hall_mu_vs_oracle(gpu, N, field, direction):
    """Moments through one or two GEMM blocks of 256 rows."""
    s = _sample(L=6, field=field)
    H, (ip, ix, dt, dx, dy) = _dxy_ham(s)
    psi = states.random_states(s.num_orb, states.sample_seeds(17, 2))
    g = tbpm.hall_mu(H, psi, N, direction)
    r = oracle.orc_tbpm_kbdc(ip, ix, dt, dx, dy, s.rescale, psi, N, direction)
    assert g.shape == (2, N, N)
    assert np.max(np.abs(g - r)) < 1e-12 * max(1.0, np.max(np.abs(r)))
    H.close()


def test_hall_mu_gram_split_over_workgroups(gpu):
    """40x40 cells (3,136 orbitals): the Gram matrix's n dimension is split over ~50
    workgroups per output tile (fixed-order sum), 96 moments (a partial
    64-wide tile in both directions)."""
    s = _sample(L=40, field=12.0)
    H, (ip, ix, dt, dx, dy) = _dxy_ham(s)
    psi = states.random_states(s.num_orb, states.sample_seeds(23, 1))
    g = tbpm.hall_mu(H, psi, 96, 1)
    r = oracle.orc_tbpm_kbdc(ip, ix, dt, dx, dy, s.rescale, psi, 96, 1)
    assert np.max(np.abs(g - r)) < 1e-12 * max(1.0, np.max(np.abs(r)))
    H.close()


def test_hall_mu_2048_kernels_vs_oracle(gpu):
    """The reference Hall example's kernel count (2048: eight Gram blocks of
    256 rows, 32 full 64-wide tiles per row, the orbital dimension split in
    six) on its lattice scaled down (graphene 50x20 with t = 3 eV, B = 100 T):
    against orc_tbpm_kbdc, and the per-launch timing sees eight Gram blocks."""
    s = tb.make_graphene_sample(50, 20, t=3.0)
    s.set_magnetic_field(100.0, 0)
    s.rescale_ham(9.0)
    H, (ip, ix, dt, dx, dy) = _dxy_ham(s)
    psi = states.random_states(s.num_orb, states.sample_seeds(31, 1))
    h = H.handle
    h.set_kernel_timing(True)
    try:
        g = tbpm.hall_mu(H, psi, 2048, 1)
        kt = h.kernel_times()
    finally:
        h.set_kernel_timing(False)
    r = oracle.orc_tbpm_kbdc(ip, ix, dt, dx, dy, s.rescale, psi, 2048, 1)
    err = float(np.max(np.abs(g - r)))
    print(f"hall 2048 kernels: max |mu_gpu - mu_oracle| = {err:.3e} (max |mu| "
          f"{np.max(np.abs(r)):.3e})", flush=True)
    assert err < 1e-12 * max(1.0, np.max(np.abs(r)))
    assert kt["gram"][1] == 8
    H.close()


@needs_ref
def test_hall_mu_vs_reference_live(gpu):
    s = _sample(L=5, field=15.0)
    H, (ip, ix, dt, dx, dy) = _dxy_ham(s)
    S, seed, N = 2, 9, 32
    ref = oracle.ref_tbpm_kbdc(seed, ip, ix, dt, s.rescale, dx, dy, S, N, 1)
    psi = np.array([oracle.ref_random_state(s.num_orb, seed * (i + 1)) for i in range(S)])
    g = tbpm.hall_mu(H, psi, N, 1)
    avg = np.zeros_like(g[0])
    for i in range(S):
        avg = avg + g[i] / S
    assert np.max(np.abs(avg - ref)) < 1e-12 * max(1.0, np.max(np.abs(ref)))
    H.close()


def test_hall_mu_needs_fp64_and_currents(gpu):
    s = _sample(L=4)
    H32, _ = _dxy_ham(s, "fp32")
    psi = states.random_states(s.num_orb, states.sample_seeds(1, 1))
    with pytest.raises(tb._lib.TBGError):
        tbpm.hall_mu(H32, psi, 8)
    H32.close()
    ip, ix, dt = oracle.build_ham_csr(s.orb_eng, s.hop_i, s.hop_j, s.hop_v)
    Hn = tbpm.Hamiltonian(s.num_orb, ip, ix, dt, precision="fp64")
    with pytest.raises(tb._lib.TBGError):
        tbpm.hall_mu(Hn, psi, 8)
    Hn.close()


def test_solver_hall_mu_and_analyzer(gpu, tmp_path):
    s = _sample(L=6, field=20.0)
    S = 3
    cfg = _config(16, S)
    cfg.dckb["n_kernel"] = 48
    cfg.dckb["direction"] = 2
    cfg.dckb["energies"] = [-0.3, 0.0, 0.2]
    cfg.dckb["ne_integral"] = 256
    solver = tb.Solver(s, cfg, directory=str(tmp_path / "sim"), prefix="h", echo_details=False)
    mu = solver.calc_hall_mu()
    assert mu.shape == (48, 48) and (tmp_path / "sim" / "h.hall_mu.npy").exists()
    ip, ix, dt, dx, dy = oracle.build_ham_dxy_loop(s.orb_eng, s.hop_i, s.hop_j, s.hop_v, s.dr)
    psi = states.random_states(s.num_orb, states.sample_seeds(1337, S))
    r = oracle.orc_tbpm_kbdc(ip, ix, dt, dx, dy, s.rescale, psi, 48, 2)
    ref = np.zeros_like(r[0])
    for i in range(S):
        ref = ref + r[i] / S
    assert np.max(np.abs(mu - ref)) < 1e-12 * max(1.0, np.max(np.abs(ref)))
    en, cond = tb.Analyzer(s, cfg, dimension=2).calc_hall_cond(mu)
    pref = 16 * s.nr_orbitals / (math.pi * s.energy_range ** 2) / s.area_unit_cell
    oc = oracle.orc_cond_from_trace(mu, [-0.3, 0.0, 0.2], s.rescale, cfg.generic["beta"], 256)
    assert np.max(np.abs(cond - pref * oc)) < 1e-10 * max(1.0, np.max(np.abs(pref * oc)))


@pytest.mark.parametrize("field,depth", [(0.0, 24), (20.0, 60)])
def test_haydock_coef_vs_oracle(gpu, field, depth):
    s = _sample(L=8, field=field)
    ip, ix, dt = oracle.build_ham_csr(s.orb_eng, s.hop_i, s.hop_j, s.hop_v)
    H = tbpm.Hamiltonian(s.num_orb, ip, ix, dt, precision="fp64")
    sites = [0, 7, 30, 31]
    local = states.random_states(len(sites), states.sample_seeds(3, 2))
    psi = np.zeros((2, s.num_orb), complex)
    psi[:, sites] = local
    a, b = tbpm.haydock_coef(H, psi, depth, s.rescale)
    for k in range(2):
        ra, rb = oracle.orc_haydock_coef(ip, ix, dt, s.rescale, psi[k], depth)
        # Lanczos amplifies rounding with depth; the first levels agree to
        # rounding, the LDOS from all of them to a loose relative bound
        assert np.max(np.abs(a[k, :12] - ra[:12])) < 1e-11
        assert np.max(np.abs(b[k, :12] - rb[:12])) < 1e-11
        en = np.linspace(-8, 8, 33)
        g = tbpm.haydock_green(a[k], b[k], en, 0.1).imag
        r = tbpm.haydock_green(ra, rb, en, 0.1).imag
        assert np.max(np.abs(g - r)) < 1e-6 * np.max(np.abs(r))
    H.close()


def test_solver_ldos_haydock(gpu, tmp_path):
    s = _sample(L=8, W=0.0)
    T, S = 32, 2
    cfg = _config(T, S)
    cfg.LDOS["site_indices"] = [2, 3, 40]
    cfg.LDOS["recursion_depth"] = 20
    cfg.LDOS["delta"] = 0.05
    solver = tb.Solver(s, cfg, directory=str(tmp_path / "sim"), prefix="y", echo_details=False)
    en, ld = solver.calc_ldos_haydock()
    assert en.shape == (2 * T + 1,) and ld.shape == (2 * T + 1,)
    assert (tmp_path / "sim" / "y.ldos_haydock.npy").exists()
    ip, ix, dt = oracle.build_ham_csr(s.orb_eng, s.hop_i, s.hop_j, s.hop_v)
    local = states.random_states(3, states.sample_seeds(1337, S))
    ca, cb = np.zeros(20, complex), np.zeros(20)
    for k in range(S):
        psi = np.zeros(s.num_orb, complex)
        psi[[2, 3, 40]] = local[k]
        a, b = oracle.orc_haydock_coef(ip, ix, dt, s.rescale, psi, 20)
        ca, cb = ca + a / S, cb + b / S
    ref = np.array([-oracle.orc_green_function(e, 0.05, ca, cb).imag / math.pi for e in en])
    assert np.max(np.abs(ld - ref)) < 1e-8 * np.max(np.abs(ref))
    assert np.all(ld >= 0)     # -Im G / pi of a Jacobi matrix with delta > 0
