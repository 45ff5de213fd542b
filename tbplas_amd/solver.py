"""Drop-in ``Solver`` for the TBPM correlation functions.

Mirrors tbplas/tbpm/solver.py: the constructor's output/config side effects
(solver.py:84-153), ``calc_corr_dos`` (256-284), ``calc_corr_ldos``
(286-315), ``calc_corr_ac_cond`` (317-352), ``calc_corr_dyn_pol``
(354-391), ``calc_corr_dc_cond`` (393-441), ``calc_hall_mu`` (443-470),
``calc_quasi_eigenstates`` (472-505), ``calc_ldos_haydock`` (507-541)
and ``calc_psi_t`` (543-600) keep their signatures and return shapes.  The f2py Fortran calls are replaced by the HIP library;
initial states come from the seeded host generator (tbplas_amd.states) and
are averaged over samples exactly as the reference accumulates
``corr += corrval / n_ran_samples``.  A failed norm check raises
``NormError`` instead of stopping the process.
"""
import math
import os
import pickle
import time

import numpy as np

from . import _lib, states, tune
from .config import Config
from .sample import Sample, ac_tile_nm, auto_tile_nm
from .tbpm import (Hamiltonian, bessel_series, corr_ac, corr_dc, corr_dos, corr_dyn_pol,
                   corr_ldos, fermi_cheb_coef, hall_mu, haydock_coef, haydock_green,
                   ldos_states, psi_t, quasi_eigenstates)

__all__ = ["Solver", "H_BAR_EV"]

H_BAR_EV = 6.582119514e-16  # base/constants.py:9


class Solver:
    """TBPM solver on one GPU (MPI size 1: rank 0 is the master)."""

    def __init__(self, sample, config, enable_mpi=False, echo_details=True, **kwargs):
        if not isinstance(sample, Sample):
            sample = Sample.from_tbplas(sample)
        self._sample = sample
        self._config = config
        self._config.check_params()
        self.rank, self.size, self.is_master = 0, 1, True
        self.mpi_enabled = False
        self._echo = echo_details
        if enable_mpi:
            self.print("MPI is not used: one process drives one GPU.")
        self._output = dict()
        self.set_output(**kwargs)

    # -- MPIEnv stand-ins at size 1 (parallel.py:155-204) --------------------
    def print(self, *args, **kwargs):
        if self.is_master and self._echo:
            print(*args, **kwargs)

    def barrier(self):
        pass

    def all_average(self, data):
        return data

    # -- output (solver.py:104-153, 236-254) -----------------------------------
    def set_output(self, directory="sim_data", prefix=None):
        self._output["directory"] = directory
        if prefix is None:
            prefix = str(int(time.time()))
        self._output["prefix"] = prefix
        if self.is_master and not os.path.exists(directory):
            os.mkdir(directory)
        self.barrier()
        for key in ("corr_DOS", "corr_LDOS", "corr_AC", "corr_dyn_pol",
                    "corr_DC", "hall_mu", "qe", "ldos_haydock", "psi_t"):
            self._output[key] = f"{directory}/{prefix}.{key}"
        self.save_config()

    def save_config(self, filename="config.pkl"):
        if self.is_master:
            d, p = self._output["directory"], self._output["prefix"]
            with open(f"{d}/{p}.{filename}", "wb") as f:
                pickle.dump(self._config, f, pickle.HIGHEST_PROTOCOL)

    def _save_data(self, data, file_name, output_format="numpy"):
        if output_format not in ("numpy", "npy"):
            raise NotImplementedError(f"Unsupported format {output_format}")
        if self.is_master:
            np.save(f"{file_name}.npy", data)

    # -- parameters (solver.py:155-234) ----------------------------------------
    def _dist_sample(self):
        return self._config.generic["nr_random_samples"]

    def _get_time_step(self):
        return 2 * math.pi / self._sample.energy_range

    def _echo_time_step_fs(self, time_step):
        self.print(f"Time step for propagation: {time_step * H_BAR_EV * 1e15:7.3f} fs\n")

    def _get_bessel_series(self, time_step):
        g = self._config.generic
        return bessel_series(time_step * self._sample.rescale, g["Bessel_precision"],
                             g["Bessel_max"])

    def _get_beta_mu_re(self):
        beta_re = self._config.generic['beta'] * self._sample.rescale
        mu_re = self._config.generic['mu'] / self._sample.rescale
        return beta_re, mu_re

    # -- GPU settings ----------------------------------------------------------
    def _gpu(self):
        g = dict(Config().gpu)
        g.update(getattr(self._config, "gpu", {}))
        return g

    def _chk(self):
        g = self._config.generic
        thr = g["wfn_check_thr"]
        gpu = self._gpu()
        if gpu["precision"] == "fp32":
            thr = max(thr, gpu["fp32_norm_thr"])
        return g["wfn_check_steps"], thr

    def _states(self, num_sample):
        gpu = self._gpu()
        seeds = states.sample_seeds(self._config.generic["seed"], num_sample, self.rank)
        return states.random_states(self._sample.num_orb, seeds, gpu["state_kind"])

    def _handle(self):
        gpu = self._gpu()
        h = _lib.default_handle(gpu["device"])
        h.set_graph(gpu["use_graph"])
        return h

    @staticmethod
    def _batches(total, per):
        i = 0
        while i < total:
            yield i, min(total, i + per)
            i += per

    def _row_order(self, gpu, key, vectors=1, rule=auto_tile_nm):
        """(row order, tile offsets) of square tiles of gpu[key] nm ('auto':
        sized by `rule`, ``auto_tile_nm`` or ``ac_tile_nm``, for `vectors`
        states per row, the sample's orbitals per nm^2 in the x-y plane and,
        for a small sample, at least two tiles per compute unit) when the
        sample has orbital positions, else (None, None)."""
        pos = self._sample.orb_pos
        if pos is None:
            return None, None
        tile = gpu.get(key, 0)
        if tile == "auto":
            # orbitals per nm^2: the median over occupied 2 nm cells (edge
            # cells and a sheared cell's bounding box would understate it)
            xy = np.asarray(pos)[:, :2]
            cell = np.floor((xy - xy.min(axis=0)) / 2.0).astype(np.int64)
            key = cell[:, 1] * (int(cell[:, 0].max()) + 1) + cell[:, 0]
            density = float(np.median(np.unique(key, return_counts=True)[1])) / 4.0
            import torch
            cus = torch.cuda.get_device_properties(gpu["device"]).multi_processor_count
            tile = rule(vectors, gpu["precision"], max(density, 1.0), n_orb=len(xy), cus=cus)
        tile = float(tile or 0)
        if tile > 0:
            return self._sample.tile_order(tile, return_tiles=True)
        return None, None

    # -- correlation functions -------------------------------------------------
    def calc_corr_dos(self):
        """(nr_time_steps+1,) complex128 DOS correlation (solver.py:256-284)."""
        time_step = self._get_time_step()
        bes = self._get_bessel_series(time_step)
        gpu = self._gpu()
        h = self._handle()
        self._sample.device = gpu["device"]
        indptr, indices, data = self._sample.build_ham_csr_device()
        num_sample = self._dist_sample()
        self._echo_time_step_fs(time_step)
        per = max(1, min(64, int(gpu["max_vectors"])))
        order, tiles = self._row_order(gpu, "tile_nm", min(per, num_sample))
        H = Hamiltonian(self._sample.num_orb, indptr, indices, data,
                        precision=gpu["precision"], handle=h, order=order, tiles=tiles)
        psi0 = self._states(num_sample)
        chk, thr = self._chk()
        T = self._config.generic["nr_time_steps"]
        tune.apply(H, min(per, num_sample), bes, gpu["autotune"])
        corr = np.zeros(T + 1, dtype=np.complex128)
        for a, b in self._batches(num_sample, per):
            c = corr_dos(H, bes, psi0[a:b], T, chk, thr)
            for s in range(b - a):
                corr = corr + c[s] / num_sample
        corr = self.all_average(corr)
        self._save_data(corr, self._output["corr_DOS"])
        return corr

    def calc_corr_ldos(self):
        """(nr_time_steps+1,) complex128 LDOS correlation (solver.py:286-315):
        random states kept on config.LDOS['site_indices'] and normalised
        (tbpm.f90:151-156), propagated as for the DOS."""
        time_step = self._get_time_step()
        bes = self._get_bessel_series(time_step)
        gpu = self._gpu()
        h = self._handle()
        self._sample.device = gpu["device"]
        indptr, indices, data = self._sample.build_ham_csr_device()
        num_sample = self._dist_sample()
        self._echo_time_step_fs(time_step)
        per = max(1, min(64, int(gpu["max_vectors"])))
        order, tiles = self._row_order(gpu, "tile_nm", min(per, num_sample))
        H = Hamiltonian(self._sample.num_orb, indptr, indices, data,
                        precision=gpu["precision"], handle=h, order=order, tiles=tiles)
        psi0 = ldos_states(self._states(num_sample), self._config.LDOS["site_indices"])
        chk, thr = self._chk()
        T = self._config.generic["nr_time_steps"]
        tune.apply(H, min(per, num_sample), bes, gpu["autotune"])
        corr = np.zeros(T + 1, dtype=np.complex128)
        for a, b in self._batches(num_sample, per):
            c = corr_ldos(H, bes, psi0[a:b], T, chk, thr)
            for s in range(b - a):
                corr = corr + c[s] / num_sample
        corr = self.all_average(corr)
        self._save_data(corr, self._output["corr_LDOS"])
        return corr

    def calc_hall_mu(self):
        """(n_kernel, n_kernel) complex128 Kubo-Bastin moments mu_mn with
        config.dckb['n_kernel'] and ['direction'] (solver.py:443-470),
        averaged over the random states."""
        gpu = self._gpu()
        h = self._handle()
        self._sample.device = gpu["device"]
        indptr, indices, hop, dx, dy = self._sample.build_ham_dxy_device()
        num_sample = self._dist_sample()
        H = Hamiltonian(self._sample.num_orb, indptr, indices, hop, dx, dy,
                        cur_scale=self._sample.rescale, precision="fp64", handle=h)
        nk = int(self._config.dckb["n_kernel"])
        direction = int(self._config.dckb["direction"])
        psi0 = self._states(num_sample)
        mu = np.zeros((nk, nk), dtype=np.complex128)
        for a, b in self._batches(num_sample, 4):
            m = hall_mu(H, psi0[a:b], nk, direction)
            for s in range(b - a):
                mu = mu + m[s] / num_sample
        mu = self.all_average(mu)
        self._save_data(mu, self._output["hall_mu"])
        return mu

    def calc_quasi_eigenstates(self, save_data=False):
        """(n_energies, num_orb) float64 quasi-eigenstates at
        config.quasi_eigenstates['energies'] (solver.py:472-505)."""
        time_step = self._get_time_step()
        bes = self._get_bessel_series(time_step)
        gpu = self._gpu()
        h = self._handle()
        self._sample.device = gpu["device"]
        indptr, indices, data = self._sample.build_ham_csr_device()
        num_sample = self._dist_sample()
        self._echo_time_step_fs(time_step)
        per = max(1, min(32, int(gpu["max_vectors"])))
        order, tiles = self._row_order(gpu, "tile_nm", 2 * min(per, num_sample))
        H = Hamiltonian(self._sample.num_orb, indptr, indices, data,
                        precision=gpu["precision"], handle=h, order=order, tiles=tiles)
        energies = np.asarray(self._config.quasi_eigenstates["energies"], dtype=np.float64)
        psi0 = self._states(num_sample)
        chk, thr = self._chk()
        T = self._config.generic["nr_time_steps"]
        tune.apply(H, 2 * min(per, num_sample), bes, gpu["autotune"])
        out = np.zeros((len(energies), self._sample.num_orb), dtype=np.float64)
        for a, b in self._batches(num_sample, per):
            q = quasi_eigenstates(H, bes, psi0[a:b], T, time_step, energies, chk, thr)
            for s in range(b - a):
                out = out + q[s] / num_sample
        out = self.all_average(out)
        if save_data:
            self._save_data(out, self._output["qe"])
        return out

    def calc_ldos_haydock(self):
        """(energies (2T+1,), ldos (2T+1,)) by the Haydock recursion at
        config.LDOS['site_indices'] (solver.py:507-541): random states on
        the sites, coefficients averaged over samples (tbpm.f90:1000-1016),
        LDOS = -Im G00(E + i delta) / pi on E = 0.5 i E_range / T, i = -T..T."""
        if self.size != 1:
            raise RuntimeError("Using more than 1 mpi process is not allowed"
                               " for ldos_haydock")
        gpu = self._gpu()
        h = self._handle()
        self._sample.device = gpu["device"]
        indptr, indices, data = self._sample.build_ham_csr_device()
        H = Hamiltonian(self._sample.num_orb, indptr, indices, data, precision="fp64",
                        handle=h)
        ld = self._config.LDOS
        sites = np.asarray(ld["site_indices"], dtype=np.int64)
        num_sample = self._config.generic["nr_random_samples"]
        seeds = states.sample_seeds(self._config.generic["seed"], num_sample, self.rank)
        local = states.random_states(len(sites), seeds, gpu["state_kind"])
        psi0 = np.zeros((num_sample, self._sample.num_orb), dtype=np.complex128)
        psi0[:, sites] = local
        depth = int(ld["recursion_depth"])
        a, b = haydock_coef(H, psi0, depth, self._sample.rescale)
        coefa = np.zeros(depth, dtype=np.complex128)
        coefb = np.zeros(depth)
        for s in range(num_sample):
            coefa = coefa + a[s] / num_sample
            coefb = coefb + b[s] / num_sample
        T = self._config.generic["nr_time_steps"]
        en_range = self._sample.energy_range
        energies = np.array([0.5 * i * en_range / T for i in range(-T, T + 1)])
        ldos = -1.0 / math.pi * haydock_green(coefa, coefb, energies, ld["delta"]).imag
        self._save_data(ldos, self._output["ldos_haydock"])
        return energies, ldos

    def calc_psi_t(self, psi_0, time_log, dt_scale=1.0, save_data=False):
        """(num_time, num_orb) complex128 wave function psi_0 propagated and
        logged at the steps of time_log (solver.py:543-600): psi_0 is
        normalised, time_log de-duplicated and sorted, every entry must lie
        in [0, nr_time_steps]."""
        if self.size != 1:
            raise RuntimeError("Using more than 1 mpi process is not allowed"
                               " for wave function propagation")
        psi_0 = np.array(psi_0, dtype=np.complex128)
        psi_0 /= np.linalg.norm(psi_0)
        time_log = np.array(sorted(set(int(t) for t in time_log)), dtype=np.int64)
        T = self._config.generic["nr_time_steps"]
        for it in time_log:
            if it not in range(T + 1):
                raise ValueError(f"time {it} out of range")
        time_step = self._get_time_step() * dt_scale
        bes = self._get_bessel_series(time_step)
        gpu = self._gpu()
        h = self._handle()
        self._sample.device = gpu["device"]
        indptr, indices, data = self._sample.build_ham_csr_device()
        self._echo_time_step_fs(time_step)
        H = Hamiltonian(self._sample.num_orb, indptr, indices, data,
                        precision=gpu["precision"], handle=h)
        chk, thr = self._chk()
        out = psi_t(H, bes, psi_0, time_log, T, chk, thr)
        if save_data:
            self._save_data(out, self._output["psi_t"])
        return out

    def calc_corr_ac_cond(self):
        """(4, nr_time_steps) complex128 AC correlation (solver.py:317-352)."""
        time_step = 0.5 * self._get_time_step()
        bes = self._get_bessel_series(time_step)
        beta_re, mu_re = self._get_beta_mu_re()
        gpu = self._gpu()
        h = self._handle()
        self._sample.device = gpu["device"]
        indptr, indices, hop, dx, dy = self._sample.build_ham_dxy_device()
        num_sample = self._dist_sample()
        self._echo_time_step_fs(time_step)
        g = self._config.generic
        coef_F = fermi_cheb_coef(g["nr_Fermi_fft_steps"], beta_re, mu_re, False,
                                 g["Fermi_cheb_precision"])
        coef_omF = fermi_cheb_coef(g["nr_Fermi_fft_steps"], beta_re, mu_re, True,
                                   g["Fermi_cheb_precision"])
        per = max(1, min(21, int(gpu["ac_states"]), int(gpu["max_vectors"])))
        order, tiles = self._row_order(gpu, "ac_tile_nm", 3 * min(per, num_sample),
                                       ac_tile_nm)
        H = Hamiltonian(self._sample.num_orb, indptr, indices, hop, dx, dy,
                        cur_scale=self._sample.rescale, precision=gpu["precision"],
                        handle=h, order=order, tiles=tiles)
        psi0 = self._states(num_sample)
        chk, thr = self._chk()
        T = g["nr_time_steps"]
        tune.apply(H, 3 * min(per, num_sample), bes, gpu["autotune"])
        corr = np.zeros((4, T), dtype=np.complex128)
        for a, b in self._batches(num_sample, per):
            c = corr_ac(H, bes, coef_F, coef_omF, psi0[a:b], T, chk, thr)
            for s in range(b - a):
                corr = corr + c[s] / num_sample
        corr = self.all_average(corr)
        self._save_data(corr, self._output["corr_AC"])
        return corr

    def calc_corr_dyn_pol(self):
        """(n_q_points, nr_time_steps) float64 dynamic-polarisation
        correlation at config.dyn_pol['q_points'] (solver.py:354-391)."""
        time_step = 0.5 * self._get_time_step()
        bes = self._get_bessel_series(time_step)
        beta_re, mu_re = self._get_beta_mu_re()
        gpu = self._gpu()
        h = self._handle()
        self._sample.device = gpu["device"]
        indptr, indices, data = self._sample.build_ham_csr_device()
        num_sample = self._dist_sample()
        self._echo_time_step_fs(time_step)
        g = self._config.generic
        coef_F = fermi_cheb_coef(g["nr_Fermi_fft_steps"], beta_re, mu_re, False,
                                 g["Fermi_cheb_precision"])
        coef_omF = fermi_cheb_coef(g["nr_Fermi_fft_steps"], beta_re, mu_re, True,
                                   g["Fermi_cheb_precision"])
        order, tiles = self._row_order(gpu, "cond_tile_nm")
        H = Hamiltonian(self._sample.num_orb, indptr, indices, data,
                        precision=gpu["precision"], handle=h, order=order, tiles=tiles)
        q_points = np.asarray(self._config.dyn_pol["q_points"], dtype=np.float64)
        pos = np.asarray(self._sample.orb_pos, dtype=np.float64)
        psi0 = self._states(num_sample)
        chk, thr = self._chk()
        T = g["nr_time_steps"]
        per = max(1, min(32, int(gpu["max_vectors"])))
        tune.apply(H, 2 * min(per, num_sample), bes, gpu["autotune"])
        corr = np.zeros((len(q_points), T), dtype=np.float64)
        for a, b in self._batches(num_sample, per):
            c = corr_dyn_pol(H, bes, coef_F, coef_omF, psi0[a:b], T, q_points, pos, chk, thr)
            for s in range(b - a):
                corr = corr + c[s] / num_sample
        corr = self.all_average(corr)
        self._save_data(corr, self._output["corr_dyn_pol"])
        return corr

    def calc_corr_dc_cond(self):
        """(corr_dos (T,), corr_dc (2, 2T, T)) complex128 (solver.py:393-441)."""
        g = self._config.generic
        tnr = g['nr_time_steps']
        en_range = self._sample.energy_range
        energies_dos = np.array([0.5 * i * en_range / tnr - en_range / 2.
                                 for i in range(tnr * 2)])
        en_limit = self._config.DC_conductivity['energy_limits']
        qe_indices = np.where((energies_dos >= en_limit[0]) &
                              (energies_dos <= en_limit[1]))[0]
        time_step = self._get_time_step()
        bes = self._get_bessel_series(time_step)
        gpu = self._gpu()
        h = self._handle()
        self._sample.device = gpu["device"]
        indptr, indices, hop, dx, dy = self._sample.build_ham_dxy_device()
        num_sample = self._dist_sample()
        self._echo_time_step_fs(time_step)
        per = max(1, min(32, int(gpu["dc_states"]), int(gpu["max_vectors"])))
        order, tiles = self._row_order(gpu, "dc_tile_nm", 2 * min(per, num_sample))
        H = Hamiltonian(self._sample.num_orb, indptr, indices, hop, dx, dy,
                        cur_scale=self._sample.rescale, precision=gpu["precision"],
                        handle=h, order=order, tiles=tiles)
        psi0 = self._states(num_sample)
        chk, thr = self._chk()
        tune.apply(H, 2 * min(per, num_sample), bes, gpu["autotune"])
        corr_dos_ = np.zeros(tnr, dtype=np.complex128)
        corr_dc_ = np.zeros((2, len(energies_dos), tnr), dtype=np.complex128)
        n_en = len(qe_indices)
        for a, b in self._batches(num_sample, per):
            d, c = corr_dc(H, bes, psi0[a:b], tnr, time_step, energies_dos,
                           qe_indices, chk, thr)
            for s in range(b - a):
                corr_dos_ = corr_dos_ + d[s] / num_sample
                corr_dc_[:, :n_en, :] = corr_dc_[:, :n_en, :] + c[s] / num_sample
        corr_dos_ = self.all_average(corr_dos_)
        corr_dc_ = self.all_average(corr_dc_)
        self._save_data(corr_dos_, self._output["corr_DOS"])
        self._save_data(corr_dc_, self._output["corr_DC"])
        return corr_dos_, corr_dc_
