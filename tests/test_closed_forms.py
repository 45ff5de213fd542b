"""Closed-form pins of the host-side restatements and of the product's
Analyzer, chosen so that a plausible mistake (a sign, an index, a dropped
term, x and y exchanged) fails:

* Peierls phases (set_mag_field, sample.pyx:18-72): the flux through every
  hexagon of an open graphene flake, and gauge invariance of the spectrum.
* The rescale estimate (get_rescale, sample.pyx:75-114): a dense row sum
  with signed on-site energies and complex hoppings.
* Analyzer.calc_ac_cond (analysis.py:161-243): a damped two-level
  correlation under a geometric window, whose windowed sine sums are
  geometric series, and the imaginary part as the circular convolution
  with the discrete Hilbert kernel (2/N) cot(pi n / N) (odd n).
* Analyzer.calc_dc_cond (analysis.py:375-458): a damped oscillation
  detuned from each quasi-eigenstate energy, whose windowed cosine sum is a
  geometric series.
* Analyzer.calc_diff_coeff (analysis.py:460-511): a constant correlation
  integrates to a line.
"""
import math

import numpy as np
import pytest

import oracle
import tbplas_amd as tb

PHI0 = 4135.666734          # h/e in T nm^2, the constant of sample.pyx:48


# -- Peierls phases ------------------------------------------------------------
def _flake(L):
    """Open L x L graphene flake: hop list with dr = pos_j - pos_i."""
    hi, hj, hv, dr, pos = oracle.graphene_hops_loop(L, L, pbc=(False, False))
    np.testing.assert_allclose(dr, pos[hj] - pos[hi], atol=1e-14)
    return hi, hj, hv, dr, pos


def _hexagons(hi, hj, n, pos):
    """All 6-cycles of the hop graph (the hexagons of a honeycomb flake),
    each ordered counter-clockwise in the x-y plane."""
    nb = [set() for _ in range(n)]
    for a, b in zip(hi.tolist(), hj.tolist()):
        nb[a].add(b)
        nb[b].add(a)
    seen, out = set(), []
    for s in range(n):
        stack = [(s, (s,))]
        while stack:
            v, path = stack.pop()
            if len(path) == 6:
                if s in nb[v] and frozenset(path) not in seen:
                    seen.add(frozenset(path))
                    out.append(list(path))
                continue
            for w in nb[v]:
                if w > s and w not in path:
                    stack.append((w, path + (w,)))
    res = []
    for c in out:
        x, y = pos[c, 0], pos[c, 1]
        area = 0.5 * np.sum(x * np.roll(y, -1) - np.roll(x, -1) * y)
        res.append((c if area > 0 else c[::-1], abs(area)))
    return res


def _amp(hi, hj, v):
    m = {}
    for a, b, z in zip(hi.tolist(), hj.tolist(), v.tolist()):
        m[(a, b)] = z
        m[(b, a)] = np.conj(z)
    return m


def _dense(n, hi, hj, v, onsite=None):
    H = np.zeros((n, n), dtype=complex)
    np.add.at(H, (hi, hj), v)
    np.add.at(H, (hj, hi), np.conj(v))
    if onsite is not None:
        H += np.diag(onsite)
    return H


@pytest.mark.parametrize("B", [20.0, 3000.0])
def test_peierls_flux_through_every_hexagon(B):
    """Gauges 0 and 1 thread 2 pi B A / Phi_0 through each hexagon (loop
    product of the amplitudes H[i_{k+1}, i_k], counter-clockwise); gauge 2
    of the reference, phase ~ dx (x_i + x_j) - dy (y_i + y_j), is an exact
    differential d(x^2 - y^2) and threads no flux at all.  The product's
    host Peierls code follows the same formulas (it is compared with this
    restatement bit for bit in the GPU builder tests)."""
    hi, hj, hv, dr, pos = _flake(5)
    hexes = _hexagons(hi, hj, len(pos), pos)
    assert len(hexes) == 16                      # (L-1)^2 interior hexagons of a 5x5 flake
    a_hex = math.sqrt(3) / 2 * 0.246 ** 2
    for gauge, flux in ((0, 1.0), (1, 1.0), (2, 0.0)):
        v = oracle.set_mag_field_loop(hi, hj, hv, dr, pos, B, gauge)
        amp = _amp(hi, hj, v)
        for cyc, area in hexes:
            assert area == pytest.approx(a_hex, rel=1e-12)
            p = 1.0 + 0j
            for k in range(6):
                p *= amp[(cyc[(k + 1) % 6], cyc[k])] / abs(amp[(cyc[(k + 1) % 6], cyc[k])])
            want = np.exp(1j * flux * 2 * math.pi * B * area / PHI0)
            assert abs(p - want) < 1e-12, (gauge, cyc)


def test_peierls_spectrum_gauge_invariant():
    """Gauges 0 and 1 are related by a diagonal unitary on an open flake:
    same spectrum, different from the field-free one; gauge 2 (no flux)
    keeps the field-free spectrum."""
    hi, hj, hv, dr, pos = _flake(4)
    n = len(pos)
    B = 4000.0
    ev = {g: np.linalg.eigvalsh(_dense(n, hi, hj, oracle.set_mag_field_loop(
        hi, hj, hv, dr, pos, B, g))) for g in (0, 1, 2)}
    e0 = np.linalg.eigvalsh(_dense(n, hi, hj, hv))
    assert np.max(np.abs(ev[0] - ev[1])) < 1e-12
    assert np.max(np.abs(ev[0] - e0)) > 1e-3
    assert np.max(np.abs(ev[2] - e0)) < 1e-12


# -- rescale estimate -----------------------------------------------------------
def test_get_rescale_closed_form_signed_onsite():
    """max_i(orb_eng_i + sum_{j != i} |H_ij|) with signed on-site energies,
    complex hoppings of varied modulus; the largest |H| row is not the
    answer when its on-site energy is negative."""
    hi, hj, hv, dr, pos = oracle.graphene_hops_loop(5, 5)
    n = len(pos)
    rng = np.random.default_rng(17)
    hv = hv * rng.uniform(0.5, 1.5, len(hv)) * np.exp(1j * rng.uniform(-3, 3, len(hv)))
    orb = rng.uniform(-2.0, 2.0, n)
    H = _dense(n, hi, hj, hv)
    want = np.max(orb + np.sum(np.abs(H), axis=1))
    got = oracle.get_rescale_loop(orb, hi, hj, hv)
    assert got == pytest.approx(want, rel=1e-14)
    s = tb.Sample(orb_eng=orb, orb_pos=pos, hop_i=hi, hop_j=hj, hop_v=hv, dr=dr)
    assert oracle.get_rescale_loop(s.orb_eng, s.hop_i, s.hop_j, s.hop_v) == got
    # two orbitals: on-site -5 and 0.1, one hop of modulus 1: rows -4 and 1.1
    got2 = oracle.get_rescale_loop(np.array([-5.0, 0.1]), np.array([0]), np.array([1]),
                                   np.array([0.6 + 0.8j]))
    assert got2 == pytest.approx(1.1, abs=1e-15)
    # the modulus is sqrt(re^2 + im^2) (sample.pyx:110): its squares underflow
    # where hypot (Python's abs) would not
    z = np.array([3e-170 + 4e-170j])
    assert abs(z[0]) == 5e-170
    assert oracle.get_rescale_loop(np.zeros(2), np.array([0]), np.array([1]), z) == 0.0


# -- Analyzer: AC ---------------------------------------------------------------
class _S:
    energy_range = 18.0
    nr_orbitals = 2
    area_unit_cell = 0.0524
    volume_unit_cell = 0.0524
    extended = 1


def _geom_cos(rho, x, T):
    """sum_{k<T} rho^k cos(k x) (geometric series)."""
    q = rho * np.exp(1j * x)
    return np.real((1 - q ** T) / (1 - q))


def _hilbert_kernel_imag(sigma):
    """Imaginary part of the discrete analytic signal of real sigma (even
    length N) as the circular convolution with h[n] = (2/N) cot(pi n/N)
    for odd n, 0 for even n."""
    N = len(sigma)
    n = np.arange(N)
    h = np.zeros(N)
    odd = n % 2 == 1
    h[odd] = 2.0 / N / np.tan(math.pi * n[odd] / N)
    idx = (n[:, None] - n[None, :]) % N
    return h[idx] @ sigma


def _ac_closed_form(T, en_range, beta, pref, r, Om, gam, amp):
    tau = math.pi / en_range
    om = np.array([i * en_range / T for i in range(T)])
    rho = r * math.exp(-gam * tau)
    # Im C_k = -amp sin(Om k tau) e^{-gam k tau}; 2 sin a sin b = cos(a-b) - cos(a+b)
    acv = -amp * r * (_geom_cos(rho, (om - Om) * tau, T) - _geom_cos(rho, (om + Om) * tau, T))
    with np.errstate(divide="ignore", invalid="ignore"):
        fac = np.where(om == 0, 0.0, (np.exp(-beta * om) - 1) / np.where(om == 0, 1, om))
    re = pref * tau * acv * fac
    re[0] = 0.0
    sigma = np.zeros(2 * T)
    sigma[T:] = re
    sigma[T - np.arange(1, T)] = re[1:]
    im = _hilbert_kernel_imag(sigma)[T:]
    return om, re + 1j * im


def _ac_input(T, en_range, Om, gam, amps):
    tau = math.pi / en_range
    k = np.arange(T)
    c = np.array([a * np.exp(-1j * Om * k * tau - gam * k * tau) for a in amps])
    c[:, 0] = 0.0              # the reference's t=1 column is zero (tbpm.f90:283-302)
    return c


def test_analyzer_ac_closed_form_oracle_and_product():
    T, en_range = 128, 18.0
    r, Om, gam = 0.97, 2.3, 0.15
    amps = [1.0, -0.4, 0.25, 2.0]
    beta = 1.0 / (8.617333262e-5 * 300)
    pref = 2 / (0.0524 * 1)
    corr = _ac_input(T, en_range, Om, gam, amps)

    def win(i, tnr):
        return r ** i

    om_o, ac_o = oracle.analyzer_calc_ac_cond(corr, T, en_range, beta, pref, window=win)
    cfg = tb.Config()
    cfg.generic["nr_time_steps"] = T
    om_p, ac_p = tb.Analyzer(_S(), cfg, dimension=2).calc_ac_cond(corr, window=win)
    for j, a in enumerate(amps):
        # k = 0 carries weight sin(0) = 0, so the zeroed column changes nothing
        om, want = _ac_closed_form(T, en_range, beta, pref, r, Om, gam, a)
        scale = np.max(np.abs(want))
        np.testing.assert_allclose(om_o, om, rtol=0, atol=1e-14)
        np.testing.assert_allclose(om_p, om, rtol=0, atol=1e-14)
        assert np.max(np.abs(ac_o[j] - want)) < 1e-11 * scale, j
        assert np.max(np.abs(ac_p[j] - want)) < 1e-11 * scale, j
    # the peak of the real part sits at the level spacing
    assert abs(om_p[np.argmax(np.abs(ac_p[0].real))] - Om) <= en_range / T


# -- Analyzer: DC ---------------------------------------------------------------
def test_analyzer_dc_closed_form_oracle_and_product():
    T, en_range = 64, 18.0
    tau = 2 * math.pi / en_range
    r, decay, detune = 0.95, 20.0, 0.3
    lim = (-1.0, 1.0)
    energies = np.array([0.5 * i * en_range / T - en_range / 2 for i in range(2 * T)])
    qe = np.where((energies >= lim[0]) & (energies <= lim[1]))[0]
    ne = len(qe)
    k = np.arange(T)
    corr_dos = np.exp(-1j * 0.7 * tau * np.arange(1, T + 1)) * np.exp(-np.arange(1, T + 1) / 40)
    corr_dc = np.zeros((2, 2 * T, T), dtype=complex)
    for a in range(2):
        for j in range(ne):
            E1 = energies[qe[j]] + detune
            corr_dc[a, j] = (a + 1) * np.exp(1j * E1 * k * tau - k / decay)
    pref = 2 / 0.0524

    def win(i, tnr):
        return r ** i

    q = r * np.exp(1j * detune * tau - 1 / decay)
    s = r * np.real((1 - q ** T) / (1 - q))
    _, dos = oracle.analyzer_calc_dos(np.insert(corr_dos, 0, 1.0), T, en_range)
    want = np.array([[pref * tau * dos[qe[j]] * (a + 1) * s for j in range(ne)]
                     for a in range(2)])
    e_o, dc_o = oracle.analyzer_calc_dc_cond(corr_dos, corr_dc, T, en_range, lim, pref,
                                             window_dc=win)
    cfg = tb.Config()
    cfg.generic["nr_time_steps"] = T
    cfg.DC_conductivity["energy_limits"] = lim
    e_p, dc_p = tb.Analyzer(_S(), cfg, dimension=2).calc_dc_cond(corr_dos, corr_dc,
                                                                   window_dc=win)
    np.testing.assert_array_equal(e_o, energies[qe])
    np.testing.assert_array_equal(e_p, energies[qe])
    assert np.max(np.abs(dc_o - want)) < 1e-12 * np.max(np.abs(want))
    assert np.max(np.abs(dc_p - want)) < 1e-12 * np.max(np.abs(want))


# -- Analyzer: diffusion coefficient ---------------------------------------------
def test_analyzer_diff_coeff_closed_form_oracle_and_product():
    """A correlation equal to c e^{i E t dt} at each energy E is the constant
    c after the phase shift: D(t) = c t with a rectangle window, and
    c dt (1 + r)/2 * r (1 - r^k)/(1 - r) under the geometric window r^i."""
    T, en_range = 48, 18.0
    tau = 2 * math.pi / en_range
    lim = (-0.8, 0.8)
    energies = np.array([0.5 * i * en_range / T - en_range / 2 for i in range(2 * T)])
    qe = np.where((energies >= lim[0]) & (energies <= lim[1]))[0]
    ne = len(qe)
    k = np.arange(T)
    cs = np.array([[0.7, -1.3, 2.0, 0.4, 1.1][j % 5] * (a + 1) for a in range(2)
                   for j in range(ne)]).reshape(2, ne)
    corr_dc = np.zeros((2, 2 * T, T), dtype=complex)
    for a in range(2):
        for j in range(ne):
            corr_dc[a, j] = cs[a, j] * np.exp(1j * energies[qe[j]] * k * tau)
    cfg = tb.Config()
    cfg.generic["nr_time_steps"] = T
    cfg.DC_conductivity["energy_limits"] = lim
    an = tb.Analyzer(_S(), cfg, dimension=2)
    r = 0.9
    for win, want_k in ((lambda i, tnr: np.ones_like(np.asarray(i, dtype=float)),
                         lambda kk: kk * tau),
                        (lambda i, tnr: r ** i,
                         lambda kk: tau * (1 + r) / 2 * r * (1 - r ** kk) / (1 - r))):
        want = cs[:, :, None] * want_k(k)[None, None, :]
        t_o, d_o = oracle.analyzer_calc_diff_coeff(corr_dc, T, en_range, lim, window_dc=win)
        t_p, d_p = an.calc_diff_coeff(corr_dc, window_dc=win)
        np.testing.assert_allclose(t_o, k * tau, rtol=1e-15, atol=0)
        np.testing.assert_array_equal(t_p, t_o)
        assert d_o.shape == d_p.shape == (2, ne, T) and d_p.dtype == np.float64
        assert np.max(np.abs(d_o - want)) < 1e-12 * np.max(np.abs(want))
        assert np.max(np.abs(d_p - want)) < 1e-12 * np.max(np.abs(want))
