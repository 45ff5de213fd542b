"""The oracle is pinned before it is trusted: against the reference's own
Fortran (compiled from its sources into oracle/_ref), the golden fixtures it
produced, and closed forms (k-space exact propagation, dense
eigendecomposition, Bessel identities, Fermi-series identities)."""
import math

import numpy as np
import pytest

import oracle
import tbplas_amd as tb
from conftest import golden

KB = 8.617333262e-5
needs_ref = pytest.mark.skipif(not oracle.ref_available(),
                               reason="oracle/_ref not built (reference sources absent)")


def _graphene(L, rescale=9.0):
    hi, hj, hv, dr, pos = oracle.graphene_hops_loop(L, L)
    n = 2 * L * L
    return np.zeros(n), hi, hj, hv * (1.0 / rescale), dr, pos


# -- C restatement vs golden outputs of the reference --------------------------
def test_orc_dos_matches_reference_fixture():
    g = golden("ref_dos.npz")
    c = oracle.orc_tbpm_dos(g["indptr"], g["indices"], g["data"], g["bes"], g["psi0"],
                            int(g["T"]))[0]
    assert np.max(np.abs(c - g["corr"])) < 1e-13


def test_orc_ac_matches_reference_fixture():
    g = golden("ref_ac.npz")
    c = oracle.orc_tbpm_accond(g["indptr"], g["indices"], g["hop"], g["dx"], g["dy"],
                               float(g["rescale"]), float(g["beta_re"]), float(g["mu_re"]),
                               g["bes"], g["psi0"], int(g["T"]))
    S = c.shape[0]
    avg = np.zeros_like(c[0])
    for s in range(S):
        avg = avg + c[s] / S
    ref = g["corr"]
    assert np.all(avg[:, 0] == 0) and np.all(ref[:, 0] == 0)     # t = 1 never stored
    assert np.max(np.abs(avg - ref)) < 1e-12 * max(1.0, np.max(np.abs(ref)))


def test_orc_dc_matches_reference_fixture():
    g = golden("ref_dc.npz")
    dos, dc = oracle.orc_tbpm_dccond(g["indptr"], g["indices"], g["hop"], g["dx"], g["dy"],
                                     float(g["rescale"]), g["bes"], g["psi0"], int(g["T"]),
                                     float(g["t_step"]), g["energies"], g["en_inds"])
    S = dos.shape[0]
    ne = len(g["en_inds"])
    adc = np.zeros_like(dos[0])
    acc = np.zeros((2, len(g["energies"]), int(g["T"])), dtype=complex)
    for s in range(S):
        adc = adc + dos[s] / S
        acc[:, :ne, :] = acc[:, :ne, :] + dc[s] / S
    assert np.max(np.abs(adc - g["dos"])) < 1e-13
    scale = max(1.0, np.max(np.abs(g["dc"])))
    assert np.max(np.abs(acc - g["dc"])) < 1e-12 * scale
    assert np.all(g["dc"][:, ne:, :] == 0)


@needs_ref
def test_orc_dos_matches_reference_live():
    orb, hi, hj, hv, dr, pos = _graphene(10)
    ip, ix, dt = oracle.build_ham_csr(orb, hi, hj, hv)
    bes = oracle.bessel_series(math.pi, 1e-14)
    n = len(orb)
    S, T, seed = 3, 40, 7
    ref = oracle.ref_tbpm_dos(bes, ip, ix, dt, seed, T, S)
    psi = np.array([oracle.ref_random_state(n, seed * (s + 1)) for s in range(S)])
    c = oracle.orc_tbpm_dos(ip, ix, dt, bes, psi, T)
    avg = np.zeros(T + 1, dtype=complex)
    for s in range(S):
        avg = avg + c[s] / S
    assert np.max(np.abs(avg - ref)) < 1e-13


@needs_ref
def test_fermi_coef_oracle_matches_reference():
    for beta, mu, om in ((350.0, 0.0, False), (350.0, 0.0, True), (1000.0, 0.1, False)):
        a = oracle.orc_fermi_coef(2 ** 14, beta, mu, om, 1e-12)
        b = oracle.ref_fermi_coef(2 ** 14, beta, mu, om, 1e-12)
        assert len(a) == len(b)
        assert np.max(np.abs(a - b)) < 1e-14


# -- closed forms ----------------------------------------------------------------
def test_orc_dos_vs_kspace_exact():
    L = 16
    orb, hi, hj, hv, dr, pos = _graphene(L)
    ip, ix, dt = oracle.build_ham_csr(orb, hi, hj, hv)
    tau = 2 * math.pi / 18.0 * 9.0
    bes = oracle.bessel_series(tau, 1e-15)
    rng = np.random.default_rng(3)
    psi = np.exp(2j * np.pi * rng.random(len(orb)))
    psi /= np.linalg.norm(psi)
    T = 100
    c = oracle.orc_tbpm_dos(ip, ix, dt, bes, psi, T)[0]
    k = oracle.kspace_graphene_corr(L, L, hv[0], psi, tau, range(T + 1))
    assert np.max(np.abs(c - k)) < 1e-12


def test_kspace_multi_equals_single_and_dense():
    """The several-state k-space oracle (used at 8.4M orbitals) equals the
    one-state form and the dense eigendecomposition, forward and backward,
    on a non-square lattice (L1 != L2 catches a transposed k grid)."""
    L1, L2 = 6, 4
    hi, hj, hv, dr, pos = oracle.graphene_hops_vec(L1, L2)
    ip, ix, dt = oracle.build_ham_csr(np.zeros(len(pos)), hi, hj, hv * (1 / 9.0))
    tau = math.pi
    rng = np.random.default_rng(8)
    psi = np.exp(2j * np.pi * rng.random((3, len(pos))))
    psi /= np.linalg.norm(psi, axis=1, keepdims=True)
    steps = [0, 1, 5, 40]
    for fwd in (True, False):
        m = oracle.kspace_graphene_corr_multi(L1, L2, hv[0] / 9.0, psi, tau, steps, fwd)
        for s in range(3):
            one = oracle.kspace_graphene_corr(L1, L2, hv[0] / 9.0, psi[s], tau, steps, fwd)
            assert np.max(np.abs(m[s] - one)) < 1e-13
            d = oracle.dense_corr(ip, ix, dt, psi[s], tau if fwd else -tau, steps)
            assert np.max(np.abs(m[s] - d)) < 1e-12


def test_orc_dos_vs_dense_eig_disordered():
    L = 6
    orb, hi, hj, hv, dr, pos = _graphene(L)
    rng = np.random.default_rng(11)
    orb = rng.uniform(-0.05, 0.05, len(orb))
    hv = hv * np.exp(1j * rng.uniform(-1, 1, len(hv)))      # complex hops
    ip, ix, dt = oracle.build_ham_csr(orb, hi, hj, hv)
    tau = math.pi
    bes = oracle.bessel_series(tau, 1e-15)
    psi = np.exp(2j * np.pi * rng.random(len(orb)))
    psi /= np.linalg.norm(psi)
    T = 50
    c = oracle.orc_tbpm_dos(ip, ix, dt, bes, psi, T)[0]
    e = oracle.dense_corr(ip, ix, dt, psi, tau, range(T + 1))
    assert np.max(np.abs(c - e)) < 1e-12
    # backward step is the inverse of the forward one
    f = oracle.orc_step(ip, ix, dt, bes, psi, True)
    b = oracle.orc_step(ip, ix, dt, bes, f, False)
    assert np.max(np.abs(b - psi)) < 1e-13


def test_bessel_fixture_and_identities():
    g = golden("bessel.npz")
    for key, x in (("pi_1e14", float(g["tau_pi"])), ("pi_1e15", float(g["tau_pi"])),
                   ("half_1e14", float(g["tau_half"]))):
        j = np.asarray(g[key])
        np.testing.assert_array_equal(j, oracle.bessel_series(x, 1e-15 if "15" in key else 1e-14))
        assert abs(j[0] + 2 * np.sum(j[2::2]) - 1.0) < 1e-14          # sum_k J_k = 1
        assert abs(j[0] ** 2 + 2 * np.sum(j[1:] ** 2) - 1.0) < 1e-14  # sum_k J_k^2 = 1
    assert len(g["pi_1e14"]) == 20 and len(g["pi_1e15"]) == 21 and len(g["half_1e14"]) == 16


def test_fermi_fixture_identities():
    g = golden("fermi.npz")
    cF, cO = np.asarray(g["coef_F"]), np.asarray(g["coef_omF"])
    assert len(cF) == 2555 and len(cO) == 2555
    n = min(len(cF), len(cO))
    # F + (1 - F) = 1: the series add to the constant 1 (c_0 = 1, others 0)
    s = cF[:n] + cO[:n]
    assert abs(s[0] - 1.0) < 1e-12 and np.max(np.abs(s[1:])) < 1e-12
    # the damped series approximates the Fermi function at x = cos(theta)
    beta = float(g["beta_re"])
    th = np.linspace(0.05, np.pi - 0.05, 33)
    x = np.cos(th)
    series = np.array([np.sum(cF * np.cos(np.arange(len(cF)) * t)) for t in th])
    fermi = 1 / (1 + np.exp(beta * x))
    assert np.max(np.abs(series - fermi)[np.abs(x) > 0.05]) < 1e-3


def test_fermi_oracle_fixture():
    g = golden("fermi.npz")
    a = oracle.orc_fermi_coef(int(g["nr_fermi"]), float(g["beta_re"]), 0.0, False, 1e-12)
    assert len(a) == len(g["coef_F"]) and np.max(np.abs(a - g["coef_F"])) < 1e-14


# -- CSR builders: loop restatement vs vectorised, fixtures, scipy --------------
def test_graphene4_fixture_hand_checked():
    g = golden("graphene4.npz")
    hi, hj = g["hop_i"], g["hop_j"]
    # cell (0,0): A0-B1 intra, B1 -> A of cell (1,0) = orbital 8, B1 -> A of (0,1) = 2
    assert list(hi[:3]) == [0, 1, 1] and list(hj[:3]) == [1, 8, 2]
    # cell (3,3) wraps to (0,3) and (3,0)
    assert list(hi[-3:]) == [30, 31, 31] and list(hj[-3:]) == [31, 6, 24]
    ip = g["dxy_indptr"]
    assert np.all(np.diff(ip) == 3)                       # no kept diagonal, 3 bonds
    # row 0: (0,1) from hop 0, then conjugates of hops landing on A0
    assert list(g["dxy_indices"][:3]) == [1, 7, 25]
    np.testing.assert_array_equal(np.sort(g["dxy_indices"][:3]), g["csr_indices"][:3])


def test_dxy_loop_vs_stable_and_scipy():
    rng = np.random.default_rng(5)
    for L, pbc in ((5, (True, True)), (4, (False, False)), (3, (True, False)), (2, (True, True))):
        hi, hj, hv, dr, pos = oracle.graphene_hops_loop(L, L, pbc=pbc)
        n = 2 * L * L
        orb = rng.uniform(-1e-4, 1e-4, n)
        orb[::3] = 0.0
        hv = hv * np.exp(1j * rng.uniform(-1, 1, len(hv)))
        a = oracle.build_ham_dxy_loop(orb, hi, hj, hv, dr)
        b = oracle.build_ham_dxy_stable(orb, hi, hj, hv, dr)
        for x, y in zip(a, b):
            np.testing.assert_array_equal(x, y)
        # sorting each row and summing duplicates reproduces scipy's canonical H
        ip, ix, dt = oracle.build_ham_csr(orb, hi, hj, hv)
        dense_a = np.zeros((n, n), complex)
        for r in range(n):
            for p in range(a[0][r], a[0][r + 1]):
                dense_a[r, a[1][p]] += a[2][p]
        dense_b = np.zeros((n, n), complex)
        for r in range(n):
            for p in range(ip[r], ip[r + 1]):
                dense_b[r, ix[p]] += dt[p]
        keep = np.abs(orb) >= 1e-5
        dense_a[np.arange(n), np.arange(n)] += np.where(keep, 0, orb)
        assert np.max(np.abs(dense_a - dense_b)) < 1e-15


def test_graphene_generator_matches_loop():
    from tbplas_amd.sample import graphene_hops
    for L1, L2, pbc in ((6, 6, (True, True)), (5, 7, (True, True)), (4, 5, (False, False)),
                        (5, 4, (True, False)), (4, 6, (False, True))):
        pos, hi, hj, hv, dr = graphene_hops(L1, L2, pbc=pbc)
        a = oracle.graphene_hops_loop(L1, L2, pbc=pbc)
        np.testing.assert_array_equal(hi, a[0])
        np.testing.assert_array_equal(hj, a[1])
        np.testing.assert_array_equal(hv, a[2])
        np.testing.assert_array_equal(dr, a[3])
        np.testing.assert_array_equal(pos, a[4])


def test_vectorised_graphene_oracle_matches_loop():
    for L1, L2, pbc in ((6, 6, (True, True)), (5, 7, (True, True)), (4, 5, (False, False)),
                        (5, 4, (True, False)), (4, 6, (False, True))):
        a = oracle.graphene_hops_loop(L1, L2, pbc=pbc)
        b = oracle.graphene_hops_vec(L1, L2, pbc=pbc)
        for x, y in zip(a, b):
            np.testing.assert_array_equal(x, y)


def test_graphene_generator_vacancies_match_loop():
    """Vacancies drop the hops that touch them and renumber the orbitals
    (SuperCell(vacancies=...), super.py:485-510); checked against the loop
    restatement with the same removal done afterwards."""
    from tbplas_amd.sample import graphene_hops, random_vacancies
    for L1, L2, pbc in ((7, 5, (True, True)), (6, 6, (False, True))):
        vac = random_vacancies(2 * L1 * L2, 0.1, L1)
        pos, hi, hj, hv, dr = graphene_hops(L1, L2, pbc=pbc, vacancies=vac)
        a_hi, a_hj, a_hv, a_dr, a_pos = oracle.graphene_hops_loop(L1, L2, pbc=pbc)
        gone = np.zeros(2 * L1 * L2, dtype=bool)
        gone[vac] = True
        keep = ~gone[a_hi] & ~gone[a_hj]
        new_id = np.cumsum(~gone) - 1
        np.testing.assert_array_equal(hi, new_id[a_hi[keep]])
        np.testing.assert_array_equal(hj, new_id[a_hj[keep]])
        np.testing.assert_array_equal(hv, a_hv[keep])
        np.testing.assert_array_equal(dr, a_dr[keep])
        np.testing.assert_array_equal(pos, a_pos[~gone])


def test_rescale_and_peierls_loops_closed_form():
    orb, hi, hj, hv, dr, pos = _graphene(4, rescale=1.0)
    # pristine graphene: every row sums 3 |t|
    assert oracle.get_rescale_loop(orb, hi, hj, hv) == pytest.approx(8.1, abs=1e-12)
    v = oracle.set_mag_field_loop(hi, hj, hv, dr, pos, 0.0, 0)
    np.testing.assert_array_equal(v, hv)                   # B = 0: no phase
    v = oracle.set_mag_field_loop(hi, hj, hv, dr, pos, 20.0, 0)
    assert np.allclose(np.abs(v), np.abs(hv), rtol=0, atol=1e-15)   # phases only


def test_analyzer_loops_textbook():
    # C(t) of a single level at E0: DOS peaks at E0 and integrates to 1
    T, en_range = 256, 18.0
    dt = 2 * math.pi / en_range
    E0 = 1.5
    corr = np.exp(-1j * E0 * dt * np.arange(T + 1))
    e, d = oracle.analyzer_calc_dos(corr, T, en_range)
    assert abs(e[np.argmax(d)] - E0) <= 0.5 * en_range / T + 1e-12
    assert abs(np.sum(d) * 0.5 * en_range / T - 1.0) < 1e-12


# -- remaining Chebyshev drivers: LDOS, psi(t), quasi-eigenstates ---------------
def _dense_eig(ip, ix, dt):
    n = len(ip) - 1
    Hd = np.zeros((n, n), dtype=complex)
    for i in range(n):
        for p in range(ip[i], ip[i + 1]):
            Hd[i, ix[p]] += dt[p]
    return np.linalg.eigh(Hd)


def _small_disordered(L=5, seed=13):
    orb, hi, hj, hv, dr, pos = _graphene(L)
    rng = np.random.default_rng(seed)
    orb = rng.uniform(-0.05, 0.05, len(orb))
    hv = hv * np.exp(1j * rng.uniform(-1, 1, len(hv)))
    return oracle.build_ham_csr(orb, hi, hj, hv)


def test_orc_ldos_psi_t_vs_dense_exact():
    """LDOS correlation sum_n |<phi_n|psi0>|^2 e^{-i E_n t tau} and the
    logged frames V e^{-i E t tau} V^H psi0 (closed forms)."""
    ip, ix, dt = _small_disordered()
    n = len(ip) - 1
    tau = math.pi
    bes = oracle.bessel_series(tau, 1e-15)
    rng = np.random.default_rng(5)
    psi = np.exp(2j * np.pi * rng.random(n))
    sites = [0, 3, 7, 11]
    w0 = np.zeros(n, complex)
    w0[sites] = psi[sites]
    w0 /= np.linalg.norm(w0)
    E, V = _dense_eig(ip, ix, dt)
    c = V.conj().T @ w0
    T = 30
    ldos = oracle.orc_tbpm_ldos(ip, ix, dt, bes, w0[None, :], T)[0]
    exact = np.array([np.sum(np.abs(c) ** 2 * np.exp(-1j * E * t * tau)) for t in range(T + 1)])
    assert np.max(np.abs(ldos - exact)) < 1e-12
    tl = [0, 1, 7, 30]
    fr = oracle.orc_psi_t(ip, ix, dt, bes, w0, tl, T)
    for i, t in enumerate(tl):
        assert np.max(np.abs(fr[i] - V @ (np.exp(-1j * E * t * tau) * c))) < 1e-12


def test_orc_quasi_eigenstates_vs_dense_exact():
    """QE(E) = sum_n phi_n <phi_n|psi0> (1 + 2 sum_k W_k cos((E - E_n) k tau)),
    normalised (the spectral form of tbpm.f90:776-812)."""
    ip, ix, dt = _small_disordered(seed=17)
    n = len(ip) - 1
    tau = math.pi
    bes = oracle.bessel_series(tau, 1e-15)
    rng = np.random.default_rng(9)
    psi = np.exp(2j * np.pi * rng.random((2, n)))
    psi /= np.linalg.norm(psi, axis=1, keepdims=True)
    E, V = _dense_eig(ip, ix, dt)
    T, energies = 40, [-0.2, 0.0, 0.13]
    q = oracle.orc_quasi_eigenstates(ip, ix, dt, bes, psi, T, tau, energies)
    k = np.arange(1, T + 1)
    W = 0.5 * (1 + np.cos(np.pi * k / T))
    for s in range(2):
        c = V.conj().T @ psi[s]
        for i, e in enumerate(energies):
            f = 1 + 2 * np.sum(W[None, :] * np.cos(np.outer(e - E, k) * tau), axis=1)
            qe = V @ (c * f)
            qe /= np.linalg.norm(qe)
            assert np.max(np.abs(q[s, i] - np.abs(qe) ** 2)) < 1e-12
    assert np.allclose(q.sum(axis=2), 1.0, atol=1e-13)


@needs_ref
def test_orc_ldos_psi_t_qe_match_reference_live():
    orb, hi, hj, hv, dr, pos = _graphene(6)
    rng = np.random.default_rng(21)
    orb = rng.uniform(-0.3, 0.3, len(orb)) / 9.0
    ip, ix, dt = oracle.build_ham_csr(orb, hi, hj, hv)
    n = len(orb)
    bes = oracle.bessel_series(math.pi, 1e-14)
    S, T, seed = 2, 24, 5
    sites = [1, 4, 9, 30]
    # LDOS: the reference's own random states kept on the sites
    ref = oracle.ref_tbpm_ldos(sites, bes, ip, ix, dt, seed, T, S)
    psi = np.array([oracle.ref_random_state(n, seed * (s + 1)) for s in range(S)])
    w0 = np.zeros_like(psi)
    w0[:, sites] = psi[:, sites]
    w0 /= np.linalg.norm(w0, axis=1, keepdims=True)
    c = oracle.orc_tbpm_ldos(ip, ix, dt, bes, w0, T)
    avg = np.zeros(T + 1, dtype=complex)
    for s in range(S):
        avg = avg + c[s] / S
    assert np.max(np.abs(avg - ref)) < 1e-13
    # psi(t), log pointer quirks included (0, T and a repeated last entry)
    p0 = psi[0] / np.linalg.norm(psi[0])
    for tl in ([0, 3, T], [2, 5], [T]):
        r = oracle.ref_tbpm_psi_t(bes, ip, ix, dt, p0, tl, T)
        o = oracle.orc_psi_t(ip, ix, dt, bes, p0, tl, T)
        assert np.max(np.abs(r - o)) < 1e-13
    # quasi-eigenstates averaged over the reference's states
    t_step = math.pi / 9.0
    en = [-0.5, 0.0, 0.7]
    r = oracle.ref_tbpm_eigenstates(bes, ip, ix, dt, seed, T, S, t_step, en)
    q = oracle.orc_quasi_eigenstates(ip, ix, dt, bes, psi, T, t_step, en)
    avg = np.zeros_like(q[0])
    for s in range(S):
        avg = avg + q[s] / S
    assert r.shape == (len(en), n)
    assert np.max(np.abs(avg - r)) < 1e-13 * max(1.0, np.max(r))


def test_orc_dyn_pol_vs_dense_exact():
    """psi1 = V g(E) V^H (rho* psi0), psi2 = V f(E) V^H psi0 with the series
    f(E) = sum_k c_k cos(k arccos E) evaluated on the eigenvalues, both
    propagated by e^{-i E tau}: Im <psi2 | rho psi1> (closed form)."""
    ip, ix, dt = _small_disordered(seed=19)
    n = len(ip) - 1
    orb, hi, hj, hv, dr, pos = _graphene(5)
    tau = math.pi / 2
    bes = oracle.bessel_series(tau, 1e-15)
    beta_re = 1.0 / (KB * 600.0) * 9.0
    cF = oracle.orc_fermi_coef(2 ** 12, beta_re, 0.02, False, 1e-12)
    cO = oracle.orc_fermi_coef(2 ** 12, beta_re, 0.02, True, 1e-12)
    rng = np.random.default_rng(4)
    psi = np.exp(2j * np.pi * rng.random((2, n)))
    psi /= np.linalg.norm(psi, axis=1, keepdims=True)
    qs = [[1.3, 0.4, 0.0], [0.0, 2.1, 0.0]]
    T = 12
    c = oracle.orc_tbpm_dyn_pol(ip, ix, dt, bes, cF, cO, psi, T, qs, pos)
    E, V = _dense_eig(ip, ix, dt)
    th = np.arccos(np.clip(E, -1, 1))
    f = np.array([np.sum(cF * np.cos(np.arange(len(cF)) * x)) for x in th])
    g = np.array([np.sum(cO * np.cos(np.arange(len(cO)) * x)) for x in th])
    for iq, q in enumerate(qs):
        rho = np.exp(1j * (pos @ np.asarray(q)))
        for s in range(2):
            a1 = g * (V.conj().T @ (np.conj(rho) * psi[s]))
            a2 = f * (V.conj().T @ psi[s])
            for t in range(T):
                ph = np.exp(-1j * E * t * tau)
                p1, p2 = V @ (ph * a1), V @ (ph * a2)
                assert abs(c[s, iq, t] - np.vdot(p2, rho * p1).imag) < 1e-12


@needs_ref
def test_orc_dyn_pol_matches_reference_live():
    orb, hi, hj, hv, dr, pos = _graphene(5)
    rng = np.random.default_rng(23)
    orb = rng.uniform(-0.3, 0.3, len(orb)) / 9.0
    ip, ix, dt, dx, dy = oracle.build_ham_dxy_loop(orb, hi, hj, hv, dr)
    n = len(orb)
    tau = math.pi / 2
    bes = oracle.bessel_series(tau, 1e-14)
    beta_re, mu_re = 1.0 / (KB * 300.0) * 9.0, 0.01
    S, T, seed = 2, 10, 3
    qs = [[0.7, 0.2, 0.0], [1.5, -0.4, 0.0]]
    ref = oracle.ref_tbpm_dyn_pol(bes, beta_re, mu_re, ip, ix, dt, 9.0, dx, dy, pos, seed, T, S,
                                  qs)
    cF = oracle.orc_fermi_coef(2 ** 15, beta_re, mu_re, False, 1e-12)
    cO = oracle.orc_fermi_coef(2 ** 15, beta_re, mu_re, True, 1e-12)
    psi = np.array([oracle.ref_random_state(n, seed * (s + 1)) for s in range(S)])
    c = oracle.orc_tbpm_dyn_pol(ip, ix, dt, bes, cF, cO, psi, T, qs, pos)
    avg = np.zeros_like(c[0])
    for s in range(S):
        avg = avg + c[s] / S
    assert ref.shape == (2, T)
    assert np.max(np.abs(avg - ref)) < 1e-12 * max(1.0, np.max(np.abs(ref)))


# -- Kubo-Bastin Hall moments and conductivity -------------------------------------
def _dxy_sample(L=4, seed=29, field=0.0):
    orb, hi, hj, hv, dr, pos = _graphene(L)
    rng = np.random.default_rng(seed)
    orb = rng.uniform(-0.3, 0.3, len(orb)) / 9.0
    if field:
        hv = hv * np.exp(1j * field * rng.uniform(-1, 1, len(hv)))
    return oracle.build_ham_dxy_loop(orb, hi, hj, hv, dr)


def test_orc_kbdc_vs_dense_exact():
    """mu_ij = g_i g_j psi^H J_x^H T_i(H) J_b T_j(H) psi with T_k(H) = V
    cos(k arccos E) V^H from the eigendecomposition (closed form)."""
    ip, ix, dt, dx, dy = _dxy_sample(field=0.7)
    n = len(ip) - 1
    E, V = _dense_eig(ip, ix, dt)
    dense = lambda vals: _dense_eig_matrix(ip, ix, vals)
    Jx, Jy = dense(1j * 9.0 * dt * dx), dense(1j * 9.0 * dt * dy)
    rng = np.random.default_rng(2)
    psi = np.exp(2j * np.pi * rng.random((2, n)))
    psi /= np.linalg.norm(psi, axis=1, keepdims=True)
    N = 12
    th = np.arccos(np.clip(E, -1, 1))
    Tk = [V @ np.diag(np.cos(k * th)) @ V.conj().T for k in range(N)]
    g = oracle.jackson_kernel(N)
    for direction, Jb in ((1, Jx), (2, Jy)):
        mu = oracle.orc_tbpm_kbdc(ip, ix, dt, dx, dy, 9.0, psi, N, direction)
        for s in range(2):
            P = np.array([T @ (Jx @ psi[s]) for T in Tk])
            W = np.array([Jb @ (T @ psi[s]) for T in Tk])
            ex = np.outer(g, g) * (P.conj() @ W.T)
            assert np.max(np.abs(mu[s] - ex)) < 1e-11 * max(1.0, np.max(np.abs(ex)))


def _dense_eig_matrix(ip, ix, vals):
    n = len(ip) - 1
    M = np.zeros((n, n), dtype=complex)
    for i in range(n):
        for p in range(ip[i], ip[i + 1]):
            M[i, ix[p]] += vals[p]
    return M


def test_cond_from_trace_vectorised_vs_loops():
    """The factored Gamma_mn sums (tbplas_amd.analysis.cond_from_trace)
    against the loop restatement on random moments."""
    from tbplas_amd.analysis import cond_from_trace
    rng = np.random.default_rng(8)
    N = 9
    m = rng.normal(size=(N, N)) + 1j * rng.normal(size=(N, N))
    mu = [-0.5, 0.0, 0.3]
    a = cond_from_trace(m, mu, 9.0, 1.0 / (KB * 300), 40)
    b = oracle.orc_cond_from_trace(m, mu, 9.0, 1.0 / (KB * 300), 40)
    assert np.max(np.abs(a - b)) < 1e-10 * max(1.0, np.max(np.abs(b)))


@needs_ref
def test_orc_kbdc_and_cond_match_reference_live():
    ip, ix, dt, dx, dy = _dxy_sample(L=5, seed=31, field=0.4)
    n = len(ip) - 1
    S, seed, N = 2, 9, 16
    psi = np.array([oracle.ref_random_state(n, seed * (s + 1)) for s in range(S)])
    for direction in (1, 2):
        ref = oracle.ref_tbpm_kbdc(seed, ip, ix, dt, 9.0, dx, dy, S, N, direction)
        mu = oracle.orc_tbpm_kbdc(ip, ix, dt, dx, dy, 9.0, psi, N, direction)
        avg = np.zeros_like(mu[0])
        for s in range(S):
            avg = avg + mu[s] / S
        assert np.max(np.abs(avg - ref)) < 1e-12 * max(1.0, np.max(np.abs(ref)))
    mus = [-0.4, 0.0, 0.25]
    rc = oracle.ref_cond_from_trace(ref, mus, 9.0, 1.0 / (KB * 300), 64)
    oc = oracle.orc_cond_from_trace(ref, mus, 9.0, 1.0 / (KB * 300), 64)
    assert np.max(np.abs(rc - oc)) < 1e-10 * max(1.0, np.max(np.abs(rc)))


def test_orc_haydock_moments_and_resolvent_exact():
    """Lanczos: the tridiagonal T_d of (a, b) reproduces <psi|H^k|psi> for
    k < 2d, and the continued fraction equals ((z - T_d)^{-1})_00."""
    ip, ix, dt = _small_disordered(L=4, seed=37)
    n = len(ip) - 1
    Hd = _dense_eig_matrix(ip, ix, dt) * 9.0
    rng = np.random.default_rng(12)
    psi = np.zeros(n, complex)
    psi[[1, 5, 6]] = np.exp(2j * np.pi * rng.random(3))
    psi /= np.linalg.norm(psi)
    d = 8
    a, b = oracle.orc_haydock_coef(ip, ix, dt, 9.0, psi, d)
    Td = np.diag(a) + np.diag(b[:-1], 1) + np.diag(b[:-1], -1)
    Hk, Tk = np.eye(n), np.eye(d)
    for k in range(2 * d):
        m_h = np.vdot(psi, Hk @ psi)
        assert abs(Tk[0, 0] - m_h) < 1e-10 * max(1.0, abs(m_h)), k
        Hk, Tk = Hk @ Hd, Tk @ Td
    for e in (-3.0, 0.2, 5.5):
        g = oracle.orc_green_function(e, 0.05, a, b)
        ex = np.linalg.inv((e + 0.05j) * np.eye(d) - Td)[0, 0]
        assert abs(g - ex) < 1e-10 * abs(ex)


@needs_ref
def test_orc_haydock_matches_reference_live():
    orb, hi, hj, hv, dr, pos = _graphene(6)
    rng = np.random.default_rng(41)
    orb = rng.uniform(-0.3, 0.3, len(orb)) / 9.0
    ip, ix, dt = oracle.build_ham_csr(orb, hi, hj, hv)
    n = len(orb)
    sites, seed, S, depth, T, delta = [3, 10, 20], 11, 2, 16, 40, 0.05
    en, ref = oracle.ref_ldos_haydock(sites, delta, 18.0, ip, ix, dt, 9.0, seed, depth, T, S)
    ca = np.zeros(depth, complex)
    cb = np.zeros(depth)
    for s in range(S):
        psi = np.zeros(n, complex)
        psi[sites] = oracle.ref_random_state(len(sites), seed * (s + 1))
        a, b = oracle.orc_haydock_coef(ip, ix, dt, 9.0, psi, depth)
        ca, cb = ca + a / S, cb + b / S
    assert np.allclose(en, [0.5 * i * 18.0 / T for i in range(-T, T + 1)])
    ld = np.array([-oracle.orc_green_function(e, delta, ca, cb).imag / math.pi for e in en])
    assert np.max(np.abs(ld - ref)) < 1e-9 * max(1.0, np.max(np.abs(ref)))


# -- general supercell hop lists --------------------------------------------------
GRAPHENE_PC = dict(
    lat=np.array([[0.246, 0, 0], [0.246 * math.cos(math.pi / 3), 0.246 * math.sin(math.pi / 3), 0],
                  [0, 0, 1.0]]),
    frac=np.array([[0.0, 0.0, 0.0], [1 / 3., 1 / 3., 0.0]]),
    ind=np.array([(0, 0, 0, 0, 1), (1, 0, 0, 1, 0), (0, 1, 0, 1, 0)]),
    eng=np.full(3, -2.7 + 0j))


def _odd_cell():
    """Three orbitals, hops up to two cells away, a 3-D lattice."""
    lat = np.array([[0.31, 0.02, 0.0], [-0.05, 0.27, 0.01], [0.0, 0.03, 0.42]])
    frac = np.array([[0.0, 0.0, 0.0], [0.5, 0.1, 0.2], [0.2, 0.6, 0.7]])
    ind = np.array([(0, 0, 0, 0, 1), (0, 0, 0, 1, 2), (1, 0, 0, 0, 0), (2, -1, 0, 2, 1),
                    (0, 1, 1, 1, 0), (-1, 2, 0, 0, 2), (0, 0, 1, 2, 2), (1, 1, -1, 1, 2)])
    eng = np.array([-1.0, 0.5 + 0.2j, -0.3, 0.1j, 0.7, -0.05 + 0.01j, 0.2, -0.4])
    return lat, frac, ind, eng


def test_orc_supercell_matches_graphene_loop():
    """The general restatement on graphene's primitive cell reproduces the
    graphene-specific loop (pinned above), periodic and open."""
    g = GRAPHENE_PC
    for L1, L2, pbc in ((4, 3, (True, True, False)), (5, 4, (True, False, False)),
                        (3, 5, (False, False, False))):
        pos, hi, hj, hv, dr = oracle.orc_supercell_hops(g["lat"], g["frac"], g["ind"],
                                                        g["eng"], (L1, L2, 1), pbc)
        ri, rj, rv, rd, rp = oracle.graphene_hops_loop(L1, L2, pbc=pbc[:2])
        np.testing.assert_array_equal(pos, rp)
        np.testing.assert_array_equal(hi, ri)
        np.testing.assert_array_equal(hj, rj)
        np.testing.assert_array_equal(dr, rd)


def test_orc_supercell_geometry_invariants():
    """Every term's dr is the primitive hop vector (R + f_j - f_i) A, its
    ket is the hop's orbital in the wrapped cell, vacancies and open
    boundaries remove exactly the terms that touch them."""
    lat, frac, ind, eng = _odd_cell()
    dim, pbc = (3, 4, 2), (True, False, True)
    vac = [(0, 1, 1, 2), (2, 3, 0, 0), (1, 0, 1, 1)]
    pos, hi, hj, hv, dr = oracle.orc_supercell_hops(lat, frac, ind, eng, dim, pbc, vac)
    full = [(a, b, c, o) for a in range(3) for b in range(4) for c in range(2) for o in range(3)]
    kept = [x for x in full if x not in vac]
    assert len(pos) == len(kept)
    count = 0
    for io, (a, b, c, o) in enumerate(kept):
        for k, h in enumerate(ind):
            if h[3] != o:
                continue
            tgt = [a + h[0], b + h[1], c + h[2]]
            if any(not pbc[x] and not 0 <= tgt[x] < dim[x] for x in range(3)):
                continue
            w = tuple(tgt[x] % dim[x] for x in range(3)) + (int(h[4]),)
            if w in vac:
                continue
            count += 1
            sel = np.where((hi == io) & (hj == kept.index(w)) & (hv == eng[k]))[0]
            assert len(sel) >= 1
            exp = (np.asarray(h[:3], float) + frac[h[4]] - frac[o]) @ lat
            assert np.min(np.abs(dr[sel] - exp).max(axis=1)) < 1e-12
    assert count == len(hi)


def test_kspace_ac_matches_restatement():
    """The exact k-space AC correlations of periodic graphene (band basis,
    Fermi polynomial on the band energies) against the C restatement of
    tbpm_accond (itself pinned against the reference) on 12x12 cells, two
    states, every stored step, both currents and both states: a swapped
    current, a sign of J or a wrong time origin fails."""
    s = tb.make_graphene_sample(12)
    s.rescale_ham(9.0)
    ip, ix, hop, dx, dy = oracle.build_ham_dxy_loop(s.orb_eng, s.hop_i, s.hop_j, s.hop_v, s.dr)
    tau = math.pi / s.energy_range * s.rescale
    bes = oracle.bessel_series(tau, 1e-14)
    beta = 1.0 / (8.617333262e-5 * 300.0) * s.rescale
    mu = 0.1 / s.rescale
    cF = oracle.orc_fermi_coef(2 ** 15, beta, mu, False, 1e-12)
    cO = oracle.orc_fermi_coef(2 ** 15, beta, mu, True, 1e-12)
    psi = tb.states.random_states(s.num_orb, tb.states.sample_seeds(5, 2))
    T = 20
    r = oracle.orc_tbpm_accond(ip, ix, hop, dx, dy, s.rescale, beta, mu, bes, psi, T)
    k = oracle.kspace_graphene_ac(12, 12, s.hop_v[:3], s.dr[:3], s.rescale, cF, cO, psi, tau,
                                  range(2, T + 1))
    assert np.max(np.abs(r[:, :, 0])) == 0.0
    assert np.max(np.abs(r)) > 1e-2
    assert np.max(np.abs(r[:, :, 1:] - k)) < 1e-13
