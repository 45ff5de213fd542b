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
