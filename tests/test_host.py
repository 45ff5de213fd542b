"""Host-side logic without a GPU: states, Config, Analyzer, bench contract."""
import math
import os
import subprocess
import sys

import numpy as np
import pytest

import oracle
import tbplas_amd as tb
from conftest import ROOT


def test_states_deterministic_normalised():
    a = tb.states.random_states(1000, [1337, 2674])
    b = tb.states.random_states(1000, [1337, 2674])
    np.testing.assert_array_equal(a, b)
    assert np.allclose(np.sum(np.abs(a) ** 2, axis=1), 1.0, atol=1e-14)
    assert np.allclose(np.abs(a[0]), 1 / math.sqrt(1000), atol=1e-15)
    assert not np.allclose(a[0], a[1])
    u = tb.states.uniform01(tb.states._key(1), 200000)
    assert 0.0 <= u.min() and u.max() < 1.0 and abs(u.mean() - 0.5) < 0.01
    e = tb.states.random_state(200000, 3, kind="exp")
    assert abs(np.sum(np.abs(e) ** 2) - 1) < 1e-12
    assert tb.states.sample_seeds(1337, 3) == [1337, 2674, 4011]


def test_splitmix64_known_answers():
    # first outputs of splitmix64 seeded with 0 (Vigna's reference values)
    out = tb.states.splitmix64(0, 3)
    assert [int(x) for x in out] == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4,
                                     0x06C45D188009454F]


def test_config_defaults_and_checks(tmp_path):
    c = tb.Config()
    assert c.generic["Bessel_max"] == 250 and c.generic["nr_time_steps"] == 1024
    assert c.generic["beta"] == 1.0 / (8.617333262e-5 * 300)
    assert c.DC_conductivity["energy_limits"] == (-0.5, 0.5)
    c.set_temperature(100.0)
    assert c.generic["beta"] == pytest.approx(1 / (8.617333262e-5 * 100))
    c.generic["bogus"] = 1
    with pytest.raises(ValueError):
        c.check_params()
    del c.generic["bogus"]
    c.gpu["precision"] = "fp16"
    with pytest.raises(ValueError):
        c.check_params()


class _S:
    energy_range = 18.0
    nr_orbitals = 2
    area_unit_cell = 0.0524
    volume_unit_cell = 0.0524
    extended = 1


def test_auto_tile_size():
    """Tiles sized for the bytes of states per row (vector slots: the
    smaller of the next power of two and the next 3 * 2^j) and the terms
    per pass T (five up to 128 bytes a row, else four): the edge a whose
    LDS rows 2 rho a^2 + (2 T - 3) * 2.96 sqrt(rho) a fill 150 KB, within
    [1.5, 7.5] nm; on graphene with four terms 5.04 nm at 64 bytes a row,
    3.41 at 128, 7.35 at 32; with five 4.83 at 64, 3.21 at 128 and 7.13 at
    32."""
    rho = 38.2

    def rows(a, terms):
        return 2 * rho * a * a + (2 * terms - 3) * 2.96 * math.sqrt(rho) * a

    for vec, prec, row, terms in ((16, "fp32", 128, 5), (10, "fp32", 96, 5),
                                  (24, "fp32", 192, 4), (16, "fp64", 256, 4),
                                  (8, "fp32", 64, 5), (4, "fp32", 32, 5), (2, "fp64", 32, 5)):
        assert tb.sample.pass_terms(vec, prec) == terms
        a = tb.auto_tile_nm(vec, prec)
        assert 1.5 < a < 7.5
        assert math.isclose(rows(a, terms) * row, 150 * 1024, rel_tol=1e-9)
        # an explicit term count sizes for it
        for t in (3, 4, 5):
            a = tb.auto_tile_nm(vec, prec, terms=t)
            assert a == 7.5 or math.isclose(rows(a, t) * row, 150 * 1024, rel_tol=1e-9)
    assert abs(tb.auto_tile_nm(16, "fp32") - 3.21) < 0.01
    assert abs(tb.auto_tile_nm(16, "fp32", terms=4) - 3.41) < 0.01
    assert abs(tb.auto_tile_nm(8, "fp32") - 4.83) < 0.01
    assert abs(tb.auto_tile_nm(8, "fp32", terms=4) - 5.04) < 0.01
    assert abs(tb.auto_tile_nm(4, "fp32") - 7.13) < 0.01
    assert abs(tb.auto_tile_nm(4, "fp32", terms=4) - 7.35) < 0.01
    assert tb.auto_tile_nm(8, "fp64") == tb.auto_tile_nm(16, "fp32")
    assert tb.auto_tile_nm(24, "fp32") == tb.auto_tile_nm(24, "fp32", terms=4)
    assert tb.auto_tile_nm(1, "fp64") == 7.5            # 16 B: capped
    assert tb.auto_tile_nm(64, "fp64") == 1.5           # 1024 B, floor
    # a denser lattice takes smaller tiles
    assert tb.auto_tile_nm(16, "fp32", density=4 * rho) < tb.auto_tile_nm(16, "fp32")
    # a small sample keeps two tiles per compute unit: 131,072 orbitals on
    # 256 CUs, sqrt(131072 / rho / 512) = 2.59 nm; a large one is unchanged
    assert abs(tb.auto_tile_nm(1, "fp64", n_orb=131072, cus=256) - 2.589) < 0.001
    assert tb.auto_tile_nm(8, "fp32", n_orb=8388608, cus=256) == tb.auto_tile_nm(8, "fp32")
    assert tb.auto_tile_nm(1, "fp64", n_orb=1000, cus=256) == 1.5
    # AC: complex64 follows the rule; complex128 stays at 2.5 nm on graphene
    assert tb.ac_tile_nm(12, "fp32") == tb.auto_tile_nm(12, "fp32")
    assert abs(tb.ac_tile_nm(12, "fp32") - 3.81) < 0.01
    assert tb.ac_tile_nm(12, "fp64") == 2.5
    assert tb.ac_tile_nm(48, "fp64") == tb.auto_tile_nm(48, "fp64") < 2.5
    assert abs(tb.ac_tile_nm(12, "fp64", density=4 * rho) - 1.5) < 1e-12


def test_analyzer_dos_matches_loop_restatement():
    cfg = tb.Config()
    T = 128
    cfg.generic["nr_time_steps"] = T
    rng = np.random.default_rng(0)
    corr = np.exp(-1j * rng.uniform(0, 3, T + 1).cumsum() / 10) * np.exp(-np.arange(T + 1) / 50)
    corr[0] = 1.0
    a = tb.Analyzer(_S(), cfg, dimension=2)
    e1, d1 = a.calc_dos(corr)
    e2, d2 = oracle.analyzer_calc_dos(corr, T, 18.0)
    np.testing.assert_array_equal(e1, e2)
    assert np.max(np.abs(d1 - d2)) < 1e-12 * np.max(d2)


def test_analyzer_ac_dc_match_loop_restatement():
    cfg = tb.Config()
    T = 64
    cfg.generic["nr_time_steps"] = T
    rng = np.random.default_rng(1)
    corr_ac = rng.normal(size=(4, T)) + 1j * rng.normal(size=(4, T))
    corr_ac[:, 0] = 0
    a = tb.Analyzer(_S(), cfg, dimension=2)
    om1, ac1 = a.calc_ac_cond(corr_ac)
    pref = 2 / (0.0524 * 1)
    om2, ac2 = oracle.analyzer_calc_ac_cond(corr_ac, T, 18.0, cfg.generic["beta"], pref)
    np.testing.assert_array_equal(om1, om2)
    assert np.max(np.abs(ac1 - ac2)) < 1e-9 * np.max(np.abs(ac2))
    corr_dos = np.exp(-np.arange(T) / 30.0) + 0j
    lim = (-0.5, 0.5)
    cfg.DC_conductivity["energy_limits"] = lim
    ne = int(np.sum((np.array([0.5 * i * 18 / T - 9 for i in range(2 * T)]) >= lim[0]) &
                    (np.array([0.5 * i * 18 / T - 9 for i in range(2 * T)]) <= lim[1])))
    corr_dc = rng.normal(size=(2, 2 * T, T)) + 1j * rng.normal(size=(2, 2 * T, T))
    e1, dc1 = a.calc_dc_cond(corr_dos, corr_dc)
    e2, dc2 = oracle.analyzer_calc_dc_cond(corr_dos, corr_dc, T, 18.0, lim, pref)
    assert len(e1) == ne
    np.testing.assert_array_equal(e1, e2)
    assert np.max(np.abs(dc1 - dc2)) < 1e-10 * np.max(np.abs(dc2))


def test_fermi_coef_host_vs_oracle():
    from tbplas_amd.tbpm import fermi_cheb_coef
    for beta, mu, om in ((350.0, 0.0, False), (350.0, 0.0, True), (1200.0, -0.05, False)):
        a = fermi_cheb_coef(2 ** 14, beta, mu, om, 1e-12)
        b = oracle.orc_fermi_coef(2 ** 14, beta, mu, om, 1e-12)
        assert len(a) == len(b) and np.max(np.abs(a - b)) < 1e-14


def test_bench_refuses_multi_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2 and "one GPU" in r.stderr


def test_bench_traffic_keyed_by_sources_and_layout(tmp_path, monkeypatch):
    """A counter profile applies only to the workload and tile layout it was
    measured on and to the sources that workload's kernels are built from:
    a change in the tile pass leaves the Hall profile valid and a change in
    the Hall code leaves the propagation profiles valid."""
    import json
    sys.path.insert(0, ROOT)
    import bench
    csrc = tmp_path / "tbplas_amd" / "csrc"
    csrc.mkdir(parents=True)
    (tmp_path / "include").mkdir()
    (tmp_path / "include" / "tbplas_amd.h").write_text("header")
    for name in ("tbg_block_pass.h", "tbg_pass_f32_t4_real.hip", "tbg_hall.hip", "tbg_prop.hip",
                 "tbg_mlp.hip"):
        (csrc / name).write_text(name)
    monkeypatch.setattr(bench, "HERE", str(tmp_path))
    dos, hall = "dos_fp32_k8@5nm", "hall"
    h_dos, h_hall = bench.src_hash(dos), bench.src_hash(hall)
    (tmp_path / "profiles").mkdir()
    (tmp_path / "profiles" / "traffic.json").write_text(json.dumps({
        dos: {"k_pass": {"src_hash": h_dos, "bytes_per_launch": 5}},
        hall: {"k_gram": {"src_hash": h_hall, "bytes_per_launch": 7}}}))
    assert bench.counter_traffic(dos, "k_pass")["bytes_per_launch"] == 5
    assert bench.counter_traffic("dos_fp32_k8@3.4nm", "k_pass") is None
    assert bench.counter_traffic(dos, "k_other") is None
    (csrc / "tbg_hall.hip").write_text("changed")
    assert bench.src_hash(dos) == h_dos and bench.src_hash(hall) != h_hall
    assert bench.counter_traffic(dos, "k_pass") is not None
    assert bench.counter_traffic(hall, "k_gram") is None
    (csrc / "tbg_hall.hip").write_text("tbg_hall.hip")
    (csrc / "tbg_mlp.hip").write_text("changed")        # built into neither
    assert bench.src_hash(dos) == h_dos and bench.src_hash(hall) == h_hall
    (csrc / "tbg_pass_f32_t4_real.hip").write_text("changed")
    assert bench.src_hash(dos) != h_dos and bench.src_hash(hall) == h_hall
    (csrc / "tbg_prop.hip").write_text("changed")       # built into both
    assert bench.src_hash(hall) != h_hall


def test_product_has_no_oracle_import():
    """The product path never imports the oracle (no CPU fallback)."""
    pkg = os.path.join(ROOT, "tbplas_amd")
    for dp, _, fs in os.walk(pkg):
        for f in fs:
            if f.endswith((".py", ".hip", ".h", ".cpp")):
                txt = open(os.path.join(dp, f)).read()
                assert "import oracle" not in txt and "from oracle" not in txt, f


def test_one_hip_runtime_when_library_loads_first():
    """The host generators may load the library before anything imports
    torch (Sample.set_magnetic_field does); the process must still map a
    single HIP runtime, the one torch brings."""
    code = ("import sys; sys.path.insert(0, %r)\n"
            "from tbplas_amd import _lib\n"
            "_lib.lib()\n"
            "import torch\n"
            "maps = open('/proc/self/maps').read().split('\\n')\n"
            "libs = sorted({l.split()[-1] for l in maps if 'libamdhip64' in l})\n"
            "print(len(libs), libs)\n" % ROOT)
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    assert r.stdout.startswith("1 "), r.stdout


def test_anderson_disorder_seeded_and_bounded():
    """orb_eng += U[-W/2, W/2]: added to the existing on-site energies,
    inside the bounds, reproducible from the seed, different for another
    seed, and with the uniform distribution's mean 0 and variance W^2/12."""
    W = 1.7
    a = tb.make_graphene_sample(64)
    a.orb_eng = np.full(a.num_orb, 0.25)
    b = tb.make_graphene_sample(64)
    b.orb_eng = np.full(b.num_orb, 0.25)
    tb.add_anderson_disorder(a, W, 7)
    tb.add_anderson_disorder(b, W, 7)
    np.testing.assert_array_equal(a.orb_eng, b.orb_eng)
    d = a.orb_eng - 0.25
    assert d.min() >= -0.5 * W and d.max() <= 0.5 * W
    assert abs(d.mean()) < 0.02 and abs(d.var() - W * W / 12) < 0.02
    c = tb.make_graphene_sample(64)
    tb.add_anderson_disorder(c, W, 8)
    assert not np.allclose(c.orb_eng, d)
    e = tb.make_graphene_sample(8)
    tb.add_anderson_disorder(e, 0.0, 3)
    assert np.all(e.orb_eng == 0.0)


@pytest.mark.parametrize("axis", [0, 1])
def test_uniaxial_strain_closed_form(axis):
    """Stretching by (1 + eps) along one axis scales that component of every
    position and bond vector; a bond at angle theta to the axis gets length
    d0 sqrt((1 + eps)^2 cos^2 theta + sin^2 theta), and its hopping becomes
    t exp(-3.37 (d - d0)/d0) with d0 = a / sqrt(3) (unstrained: unchanged)."""
    eps = 0.04
    s = tb.make_graphene_sample(6)
    pos0, dr0, hv0 = s.orb_pos.copy(), s.dr.copy(), s.hop_v.copy()
    tb.apply_uniaxial_strain(s, eps, axis=axis)
    d0 = 0.246 / math.sqrt(3.0)
    other = 1 - axis
    np.testing.assert_allclose(s.orb_pos[:, axis], pos0[:, axis] * (1 + eps), rtol=1e-15)
    np.testing.assert_array_equal(s.orb_pos[:, other], pos0[:, other])
    np.testing.assert_allclose(s.dr[:, axis], dr0[:, axis] * (1 + eps), rtol=1e-15)
    np.testing.assert_array_equal(s.dr[:, other], dr0[:, other])
    cos_t = dr0[:, axis] / np.linalg.norm(dr0, axis=1)
    d = d0 * np.sqrt((1 + eps) ** 2 * cos_t ** 2 + (1 - cos_t ** 2))
    np.testing.assert_allclose(s.hop_v, -2.7 * np.exp(-3.37 * (d - d0) / d0), rtol=1e-13)
    # graphene's bonds lie at 30, 150 and 90 degrees to x: stretched along y
    # the 90-degree bond loses exp(-3.37 eps); along x it keeps -2.7
    if axis == 1:
        assert np.min(np.abs(s.hop_v)) == pytest.approx(2.7 * math.exp(-3.37 * eps), rel=1e-12)
    else:
        assert np.max(np.abs(s.hop_v)) == pytest.approx(2.7, rel=1e-12)
    u = tb.make_graphene_sample(6)
    tb.apply_uniaxial_strain(u, 0.0, axis=axis)
    np.testing.assert_allclose(u.hop_v, hv0, rtol=1e-13)
