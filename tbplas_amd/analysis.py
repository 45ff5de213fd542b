"""Drop-in ``Analyzer``: correlation functions -> DOS, AC and DC conductivity.

Same formulas, windows and output grids as tbplas/tbpm/analysis.py:20-458,
vectorised with numpy (the reference loops in Python, O(4 T^2) for AC).
"""
import math

import numpy as np

__all__ = ["window_hanning", "window_exp", "window_exp_ten", "Analyzer"]


def window_hanning(i, tnr):
    """analysis.py:20-28 (works on arrays)."""
    return 0.5 * (1 + np.cos(np.pi * np.asarray(i, dtype=float) / tnr))


def window_exp(i, tnr):
    """analysis.py:31-39"""
    return np.exp(-2. * (np.asarray(i, dtype=float) / tnr) ** 2)


def window_exp_ten(i, tnr):
    """analysis.py:42-51"""
    power = -2 * (1. * np.asarray(i, dtype=float) / tnr) ** 2
    return 10. ** power


def fermi_dist(beta, ef, energy):
    """Fermi_dist (funcs.f90:36-57), elementwise over energy."""
    energy = np.asarray(energy, dtype=np.float64)
    up = energy >= ef
    x = np.exp(np.where(up, beta * (ef - energy), beta * (energy - ef)))
    return np.where(up, x / (1 + x), 1 / (1 + x))


def cond_from_trace(mu_mn, mu, h_rescale, beta, ne_integral):
    """cond_from_trace (analysis.f90:6-69) for the (N, N) moments mu_mn and
    chemical potentials mu: on the grid x_k = k pi / NE (k = 1..NE-1),
    s_k = sum_ij Re(Gamma_ij(x_k) mu_ij) with Gamma of get_gamma_mn
    (kpm.f90:31-56), then cond = sum_k s_k f(cos x_k rescale) dx / sin^3 x_k.
    Gamma_ij = c_i A_j + B_i c_j with c_i = cos(i x), A_j = (cos x - i j
    sin x) e^{i j x}, B_i = (cos x + i i sin x) e^{-i i x}, so s_k comes
    from two matrix products with mu for all k at once."""
    mu_mn = np.asarray(mu_mn, dtype=np.complex128)
    N = mu_mn.shape[0]
    dE = math.pi / ne_integral
    x = np.arange(1, ne_integral) * dE                          # (NE,)
    idx = np.arange(N)
    cx, sx = np.cos(x)[:, None], np.sin(x)[:, None]
    ph = np.outer(x, idx)                                       # (NE, N)
    c = np.cos(ph)
    A = (cx - 1j * idx[None, :] * sx) * np.exp(1j * ph)
    B = (cx + 1j * idx[None, :] * sx) * np.exp(-1j * ph)
    MA = mu_mn @ A.T                                            # (N, NE)
    Mc = mu_mn @ c.T
    s = np.real(np.sum(c.T * MA, axis=0) + np.sum(B.T * Mc, axis=0))
    en = np.cos(x) * h_rescale
    w = dE / np.sin(x) ** 3
    return np.array([np.sum(s * fermi_dist(beta, m, en) * w) for m in np.atleast_1d(mu)])


class Analyzer:
    """Post-processing on the host (analysis.py:54-87)."""

    def __init__(self, sample, config, dimension=3, enable_mpi=False, echo_details=False):
        self._sample = sample
        self._config = config
        self._config.check_params()
        if dimension not in (2, 3):
            raise ValueError(f"Unsupported dimension: {dimension}")
        self._dimension = dimension
        self.is_master = True

    def calc_dos(self, corr_dos, window=window_hanning):
        """analysis.py:89-143: (energies (2T,), dos (2T,))."""
        tnr = self._config.generic['nr_time_steps']
        en_range = self._sample.energy_range
        en_step = 0.5 * en_range / tnr
        energies = np.array([0.5 * i * en_range / tnr - en_range / 2.
                             for i in range(tnr * 2)], dtype=float)
        corr_dos = np.asarray(corr_dos)
        cn = np.empty(tnr * 2, dtype=complex)
        cn[tnr - 1] = corr_dos[0]
        cn[2 * tnr - 1] = window(tnr - 1, tnr) * corr_dos[tnr]
        i = np.arange(tnr - 1)
        w = window(i, tnr)
        cn[tnr + i] = w * corr_dos[i + 1]
        cn[tnr - i - 2] = w * np.conj(corr_dos[i + 1])
        cf = np.fft.ifft(cn)
        dos = np.empty(tnr * 2, dtype=float)
        dos[tnr:] = np.abs(cf[:tnr])
        dos[:tnr] = np.abs(cf[tnr:])
        dos = dos / (np.sum(dos) * en_step)
        if self._config.generic['correct_spin']:
            dos = 2. * dos
        return energies, dos

    def calc_ldos(self, corr_ldos, window=window_hanning):
        return self.calc_dos(corr_ldos, window)

    def _prefactor(self):
        s = self._sample
        if self._dimension == 2:
            return s.nr_orbitals / (s.area_unit_cell * s.extended), \
                s.nr_orbitals / s.area_unit_cell
        return s.nr_orbitals / (s.volume_unit_cell * s.extended), \
            s.nr_orbitals / s.volume_unit_cell

    def calc_ac_cond(self, corr_ac, window=window_exp):
        """analysis.py:161-243: (omegas (T,), ac_cond (4, T))."""
        from scipy.signal import hilbert
        tnr = self._config.generic['nr_time_steps']
        en_range = self._sample.energy_range
        t_step = math.pi / en_range
        beta = self._config.generic['beta']
        ac_prefactor = self._prefactor()[0]
        omegas = np.array([i * en_range / tnr for i in range(tnr)], dtype=float)
        k = np.arange(tnr)
        w = 2. * window(k + 1, tnr)
        sin_tab = np.sin(np.outer(omegas, k * t_step))         # (omega, k)
        corr_im = np.asarray(corr_ac).imag                      # (4, k)
        acv = (w[None, :] * corr_im) @ sin_tab.T                # (4, omega)
        with np.errstate(divide="ignore", invalid="ignore"):
            fac = np.where(omegas == 0., 0., (np.exp(-beta * omegas) - 1) / np.where(
                omegas == 0., 1., omegas))
        ac_real = ac_prefactor * t_step * acv * fac[None, :]
        ac_real[:, omegas == 0.] = 0.
        ac_imag = np.zeros((4, tnr))
        for j in range(4):
            sigma = np.zeros(2 * tnr)
            sigma[tnr:] = ac_real[j]
            sigma[tnr - np.arange(1, tnr)] = ac_real[j, 1:]
            sigma[tnr] = ac_real[j, 0]
            ac_imag[j] = np.imag(hilbert(sigma))[tnr:2 * tnr]
        ac_cond = ac_real + 1j * ac_imag
        if self._config.generic['correct_spin']:
            ac_cond = 2. * ac_cond
        return omegas, ac_cond

    def calc_hall_cond(self, mu_mn, unit="h_bar"):
        """analysis.py:513-573: (energies, Hall conductivity) at
        config.dckb['energies'] from the moments of Solver.calc_hall_mu,
        in e^2/h_bar (unit 'h_bar') or e^2/h (unit 'h')."""
        if unit not in ("h_bar", "h"):
            raise ValueError(f"Illegal unit {unit}")
        s = self._sample
        energies = np.array(self._config.dckb['energies'])
        pref = 16 * s.nr_orbitals / (math.pi * s.energy_range ** 2)
        pref /= s.area_unit_cell if self._dimension == 2 else s.volume_unit_cell
        if unit == "h":
            pref *= 2 * math.pi
        cond = cond_from_trace(mu_mn, energies, s.rescale, self._config.generic['beta'],
                               self._config.dckb['ne_integral']) * pref
        if self._config.generic['correct_spin']:
            cond = cond * 2
        return energies, cond

    def calc_dc_cond(self, corr_dos, corr_dc, window_dos=window_hanning,
                     window_dc=window_exp):
        """analysis.py:375-458: (energies (n_E,), dc (2, n_E))."""
        corr_dos = np.insert(np.asarray(corr_dos), 0, 1.0)
        energies_dos, dos = self.calc_dos(corr_dos, window_dos)
        tnr = self._config.generic['nr_time_steps']
        en_range = self._sample.energy_range
        t_step = 2 * math.pi / en_range
        lim = self._config.DC_conductivity['energy_limits']
        qe = np.where((energies_dos >= lim[0]) & (energies_dos <= lim[1]))[0]
        energies = energies_dos[qe]
        dc_prefactor = self._prefactor()[1]
        k = np.arange(tnr)
        w = window_dc(k + 1, tnr)
        phi = np.outer(energies, k * t_step)                    # (E, k)
        c_exp = np.cos(phi) - 1j * np.sin(phi)
        corr_dc = np.asarray(corr_dc)[:, :len(qe), :tnr]
        dc_val = np.sum(w[None, None, :] * (c_exp[None] * corr_dc).real, axis=2)
        dc = dc_prefactor * t_step * dos[qe][None, :] * dc_val
        if self._config.generic['correct_spin']:
            dc = 2. * dc
        return energies, dc

    def calc_diff_coeff(self, corr_dc, window_dc=window_exp):
        """analysis.py:460-511: (time (T,), diff_coeff (2, n_E, T)) -- the
        running trapezoidal integral over time of the windowed, phase-shifted
        DC correlation, in nm^2/(h_bar/eV), at the quasi-eigenstate energies
        inside DC_conductivity['energy_limits'].  Real float64, as the
        reference fills it; the reference's scipy.integrate.trapz (removed
        from scipy) becomes a cumulative sum of the same trapezoids."""
        tnr = self._config.generic['nr_time_steps']
        en_range = self._sample.energy_range
        t_step = 2 * math.pi / en_range
        lim = self._config.DC_conductivity['energy_limits']
        energies = np.array([0.5 * i * en_range / tnr - en_range / 2.
                             for i in range(tnr * 2)], dtype=float)
        qe = np.where((energies >= lim[0]) & (energies <= lim[1]))[0]
        energies = energies[qe]
        time = np.linspace(0, tnr - 1, tnr) * t_step
        k = np.arange(tnr)
        w = window_dc(k + 1, tnr)
        phi = np.outer(energies, k * t_step)                    # (E, k)
        c_exp = np.cos(phi) - 1j * np.sin(phi)
        corr_dc = np.asarray(corr_dc)[:, :len(qe), :tnr]
        temp = w[None, None, :] * (c_exp[None] * corr_dc).real   # (2, E, k)
        diff_coeff = np.zeros((2, len(qe), tnr))
        trap = 0.5 * np.diff(time)[None, None, :] * (temp[:, :, 1:] + temp[:, :, :-1])
        diff_coeff[:, :, 1:] = np.cumsum(trap, axis=2)
        return time, diff_coeff
