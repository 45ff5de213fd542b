"""The drop-in API end to end: Solver.calc_corr_* against the oracle on the
same seeded states, averaged like the reference, plus the Analyzer."""
import math
import os

import numpy as np
import pytest

import oracle
import tbplas_amd as tb
from tbplas_amd import states

pytestmark = pytest.mark.gpu


def _config(T, S, precision="fp64"):
    c = tb.Config()
    c.generic["nr_time_steps"] = T
    c.generic["nr_random_samples"] = S
    c.gpu["precision"] = precision
    c.gpu["autotune"] = False
    return c


def _avg(rows):
    out = np.zeros_like(rows[0])
    for r in rows:
        out = out + r / len(rows)
    return out


def test_calc_corr_dos_dropin(gpu, tmp_path):
    s = tb.make_graphene_sample(12)
    s.rescale_ham(9.0)
    cfg = _config(64, 3)
    solver = tb.Solver(s, cfg, directory=str(tmp_path / "sim"), prefix="t", echo_details=False)
    corr = solver.calc_corr_dos()
    assert corr.shape == (65,) and corr.dtype == np.complex128
    assert os.path.exists(tmp_path / "sim" / "t.corr_DOS.npy")
    assert os.path.exists(tmp_path / "sim" / "t.config.pkl")
    ip, ix, dt = oracle.build_ham_csr(s.orb_eng, s.hop_i, s.hop_j, s.hop_v)
    bes = oracle.bessel_series(2 * math.pi / s.energy_range * s.rescale,
                               cfg.generic["Bessel_precision"])
    psi = states.random_states(s.num_orb, states.sample_seeds(1337, 3))
    ref = _avg(oracle.orc_tbpm_dos(ip, ix, dt, bes, psi, 64))
    assert np.max(np.abs(corr - ref)) < 1e-12
    saved = np.load(tmp_path / "sim" / "t.corr_DOS.npy")
    np.testing.assert_array_equal(saved, corr)
    e, d = tb.Analyzer(s, cfg, dimension=2).calc_dos(corr)
    assert abs(np.sum(d) * 0.5 * s.energy_range / 64 - 1.0) < 1e-12
    re = tb.read_config(str(tmp_path / "sim" / "t.config.pkl"))
    assert re.generic["nr_time_steps"] == 64


def test_calc_corr_ac_dc_dropin(gpu, tmp_path):
    s = tb.make_graphene_sample(8)
    s.set_magnetic_field(30.0)
    s.rescale_ham(9.0)
    T, S = 16, 2
    cfg = _config(T, S)
    cfg.DC_conductivity["energy_limits"] = (-1.2, 1.2)
    solver = tb.Solver(s, cfg, directory=str(tmp_path / "sim"), prefix="a", echo_details=False)
    ac = solver.calc_corr_ac_cond()
    dos, dc = solver.calc_corr_dc_cond()
    assert ac.shape == (4, T) and dos.shape == (T,) and dc.shape == (2, 2 * T, T)
    ip, ix, h, dx, dy = oracle.build_ham_dxy_loop(s.orb_eng, s.hop_i, s.hop_j, s.hop_v, s.dr)
    psi = states.random_states(s.num_orb, states.sample_seeds(1337, S))
    beta_re = cfg.generic["beta"] * s.rescale
    bes_h = oracle.bessel_series(0.5 * 2 * math.pi / s.energy_range * s.rescale, 1e-14)
    ref_ac = _avg(oracle.orc_tbpm_accond(ip, ix, h, dx, dy, s.rescale, beta_re, 0.0, bes_h, psi, T))
    assert np.max(np.abs(ac - ref_ac)) < 1e-11 * max(1.0, np.max(np.abs(ref_ac)))
    en = np.array([0.5 * i * s.energy_range / T - s.energy_range / 2. for i in range(2 * T)])
    inds = np.where((en >= -1.2) & (en <= 1.2))[0]
    bes = oracle.bessel_series(2 * math.pi / s.energy_range * s.rescale, 1e-14)
    rd, rc = oracle.orc_tbpm_dccond(ip, ix, h, dx, dy, s.rescale, bes, psi, T,
                                    2 * math.pi / s.energy_range, en, inds)
    assert np.max(np.abs(dos - _avg(rd))) < 1e-12
    ne = len(inds)
    assert np.max(np.abs(dc[:, :ne, :] - _avg(rc))) < 1e-10 * max(1.0, np.max(np.abs(rc)))
    assert np.all(dc[:, ne:, :] == 0)
    an = tb.Analyzer(s, cfg, dimension=2)
    om, sig = an.calc_ac_cond(ac)
    e, cond = an.calc_dc_cond(dos, dc)
    assert sig.shape == (4, T) and cond.shape == (2, ne)


def test_solver_accepts_reference_like_sample(gpu, tmp_path):
    """A reference Sample exposes init_* methods and array attributes."""
    g = tb.make_graphene_sample(6)

    class RefLike:
        def __init__(self):
            self.orb_eng, self.orb_pos = g.orb_eng.copy(), g.orb_pos.copy()
            self.hop_i, self.hop_j, self.hop_v, self.dr = g.hop_i, g.hop_j, g.hop_v * (1 / 9.), g.dr
            self.rescale, self.area_unit_cell, self.volume_unit_cell = 9.0, 0.05, 0.05
            self.extended, self.nr_orbitals = 1, 2
            self.orb_eng = self.orb_eng * (1 / 9.)

        def init_orb_eng(self): pass
        def init_orb_pos(self): pass
        def init_hop(self): pass

    cfg = _config(8, 1, "fp32")
    solver = tb.Solver(RefLike(), cfg, directory=str(tmp_path / "s"), echo_details=False)
    c = solver.calc_corr_dos()
    assert c.shape == (9,) and abs(c[0] - 1.0) < 1e-6


def test_solver_norm_error_instead_of_stop(gpu, tmp_path):
    s = tb.make_graphene_sample(6)
    s.rescale_ham(1.5)                      # below the spectral radius: diverges
    cfg = _config(32, 1)
    cfg.generic["wfn_check_steps"] = 8
    solver = tb.Solver(s, cfg, directory=str(tmp_path / "s"), echo_details=False)
    from tbplas_amd._lib import NormError
    with pytest.raises(NormError):
        solver.calc_corr_dos()
