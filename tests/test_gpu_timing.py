"""Per-launch timing and the tile schedule's byte counts (the measurement
entry points behind bench.py's roofline records)."""
import math

import numpy as np
import pytest

import oracle
import tbplas_amd as tb
from tbplas_amd import tbpm
from conftest import graphene_sample, tau_of

pytestmark = pytest.mark.gpu


def _csr(s):
    return oracle.build_ham_csr(s.orb_eng, s.hop_i, s.hop_j, s.hop_v)


def test_kernel_times_count_every_pass_and_leave_results_unchanged(gpu):
    """21 Bessel terms with four terms per pass: per step one pass from v_0
    (mode 1), three full passes (mode 0) and one ending pass with dot and
    norm (mode 4); timing switches graph replay off and changes no bit of
    the correlation."""
    s = graphene_sample(96)
    ip, ix, dt = _csr(s)
    order, tiles = s.tile_order(2.0, return_tiles=True)
    bes = tbpm.bessel_series(tau_of(s), 1e-15)
    assert len(bes) == 21
    psi = tb.states.random_states(s.num_orb, tb.states.sample_seeds(1337, 8))
    h = tb._lib.Handle(0)
    try:
        h.set_blocking(True, terms=4)
        H = tbpm.Hamiltonian(s.num_orb, ip, ix, dt, precision="fp32", handle=h, order=order,
                             tiles=tiles)
        ref = tbpm.corr_dos(H, bes, psi, 20)
        run = tbpm.DosRun(H, bes, psi, 20)
        run.advance(4)
        h.set_kernel_timing(True)
        run.advance(12)
        kt = h.kernel_times()
        h.set_kernel_timing(False)
        run.advance(4)
        c = run.result()
        run.close()
        assert h.kernel_times() == {}
        H.close()
    finally:
        h.close()
    np.testing.assert_array_equal(c, ref)
    assert kt["pass0x4"][1] == 36 and kt["pass1x4"][1] == 12 and kt["pass4x4"][1] == 12
    assert set(kt) == {"pass0x4", "pass1x4", "pass4x4"}
    assert all(ms > 0 for ms, _ in kt.values())


def test_kernel_times_skip_the_fermi_set_up(gpu):
    """AC run: the Fermi operators (hundreds of terms, applied once) are not
    timed; every timed step has the propagation passes and one current
    correlation launch."""
    s = graphene_sample(24)
    ip, ix, hop, dx, dy = oracle.build_ham_dxy_loop(s.orb_eng, s.hop_i, s.hop_j, s.hop_v, s.dr)
    bes = tbpm.bessel_series(tau_of(s, 0.5), 1e-14)
    beta = 1.0 / (8.617333262e-5 * 300.0) * s.rescale
    cF = tbpm.fermi_cheb_coef(2 ** 12, beta, 0.0, False, 1e-12)
    cO = tbpm.fermi_cheb_coef(2 ** 12, beta, 0.0, True, 1e-12)
    assert len(cF) > 100
    psi = tb.states.random_states(s.num_orb, tb.states.sample_seeds(3, 2))
    h = tb._lib.Handle(0)
    try:
        H = tbpm.Hamiltonian(s.num_orb, ip, ix, hop, dx, dy, cur_scale=s.rescale,
                             precision="fp64", handle=h)
        T = 6
        h.set_kernel_timing(True)
        a = tbpm.corr_ac(H, bes, cF, cO, psi, T)
        kt = h.kernel_times()
        h.set_kernel_timing(False)
        b = tbpm.corr_ac(H, bes, cF, cO, psi, T)
        H.close()
    finally:
        h.close()
    np.testing.assert_array_equal(a, b)
    assert kt["ac_corr"][1] == T - 1
    assert kt["term"][1] == (T - 1) * (len(bes) - 1)


@pytest.mark.parametrize("prec,K", [("fp32", 8), ("fp64", 1), ("fp32", 24)])
def test_blocking_bytes_counts(gpu, prec, K):
    """The schedule counts agree with the tables' own summary: owned rows =
    n, rows of v_{m-2} = halo x n, at least as many distinct rows gathered
    (the ring beyond the last), and the tables hold at least one column and
    one value per entry of every row computed."""
    s = graphene_sample(64)
    ip, ix, dt = _csr(s)
    order, tiles = s.tile_order(2.0, return_tiles=True)
    h = tb._lib.Handle(0)
    try:
        h.set_blocking(True, terms=4)
        H = tbpm.Hamiltonian(s.num_orb, ip, ix, dt, precision=prec, handle=h, order=order,
                             tiles=tiles)
        b = H.blocking_bytes(K)
        nb, halo = H.blocking_info(K)
        h.set_blocking(False)
        assert H.blocking_bytes(K) is None
        H.close()
    finally:
        h.close()
    n = s.num_orb
    assert b["owned_rows"] == n
    assert b["streamed_rows"] == round(halo * n)
    assert b["streamed_rows"] < b["gathered_rows"] < 2 * b["streamed_rows"]
    vsz = 4 if prec == "fp32" else 8       # real graphene values
    assert b["tables"] >= b["streamed_rows"] * 3 * (4 + vsz)
    slots = {8: 8, 1: 1, 24: 24}[K]
    assert b["row_bytes"] == slots * (8 if prec == "fp32" else 16)
    assert b["pass_bytes"] == b["tables"] + b["row_bytes"] * (
        b["gathered_rows"] + b["streamed_rows"] + 4 * n)
    assert nb > 1 and math.isfinite(halo)
