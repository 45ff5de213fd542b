"""tbg_corr_dos_noalloc (the DOS correlation without device allocation)
against tbg_corr_dos, through the C ABI: bit-identical correlations on two
calls in a row (the second smaller, so it runs over the first one's leftover
scratch contents), the same norm-check failure, and a clear error for a run
larger than the scratch array."""
import ctypes

import numpy as np
import pytest

import tbplas_amd as tb
from tbplas_amd import _lib, tbpm
from conftest import graphene_sample, tau_of

pytestmark = pytest.mark.gpu


def _noalloc(H, bes, psi, T, chk_step=128, chk_thr=1e-9):
    torch = _lib.torch_cuda()
    dv = f"cuda:{H.handle.device}"
    p = torch.from_numpy(np.ascontiguousarray(psi, dtype=np.complex128)).to(dv)
    K, n = psi.shape
    out = torch.full((K, T + 1), complex("nan"), dtype=torch.complex128, device=dv)
    bes = np.ascontiguousarray(bes, dtype=np.float64)
    fs = ctypes.c_int(0)
    rc = _lib.lib().tbg_corr_dos_noalloc(H.handle.ptr, H.ptr, bes.ctypes.data, len(bes),
                                         _lib.ptr(p), K, int(T), int(chk_step), float(chk_thr),
                                         _lib.ptr(out), ctypes.byref(fs))
    _noalloc.msg = _lib.lib().tbg_last_error()
    return rc, fs.value, out.cpu().numpy()


def _case(L, K, T, prec, seed):
    s = tb.make_graphene_sample(L)
    tb.add_anderson_disorder(s, 0.4, seed)
    s.rescale_ham(9.0)
    ip, ix, dt = s.build_ham_csr_device()
    bes = tbpm.bessel_series(tau_of(s), 1e-15)
    psi = tb.states.random_states(s.num_orb, tb.states.sample_seeds(seed, K))
    H = tbpm.Hamiltonian(s.num_orb, ip, ix, dt, precision=prec)
    return H, bes, psi


def test_noalloc_equals_corr_dos_two_calls(gpu):
    # first call: 2 complex128 states on 64x64 graphene
    H, bes, psi = _case(64, 2, 40, "fp64", 11)
    rc, _, a = _noalloc(H, bes, psi, 40)
    assert rc == _lib.TBG_OK, _lib.lib().tbg_last_error()
    b = tbpm.corr_dos(H, bes, psi, 40)
    H.close()
    np.testing.assert_array_equal(a, b)
    assert abs(a[0, 0] - 1.0) < 1e-12
    # second call, smaller: one complex64 state on 24x24, fewer steps
    H, bes, psi = _case(24, 1, 17, "fp32", 12)
    rc, _, a = _noalloc(H, bes, psi, 17)
    assert rc == _lib.TBG_OK, _lib.lib().tbg_last_error()
    b = tbpm.corr_dos(H, bes, psi, 17)
    H.close()
    np.testing.assert_array_equal(a, b)
    # and a third, larger again: three complex64 states (three per lane)
    H, bes, psi = _case(48, 3, 25, "fp32", 13)
    rc, _, a = _noalloc(H, bes, psi, 25)
    assert rc == _lib.TBG_OK, _lib.lib().tbg_last_error()
    b = tbpm.corr_dos(H, bes, psi, 25)
    H.close()
    np.testing.assert_array_equal(a, b)


def test_noalloc_norm_check_and_size_limit(gpu):
    # rescale below the spectral radius: the norm check fails at step 16
    s = graphene_sample(8, rescale=2.0)
    ip, ix, dt = s.build_ham_csr_device()
    bes = tbpm.bessel_series(tau_of(s), 1e-14)
    psi = tb.states.random_states(s.num_orb, tb.states.sample_seeds(1, 1))
    H = tbpm.Hamiltonian(s.num_orb, ip, ix, dt, precision="fp64")
    rc, fs, _ = _noalloc(H, bes, psi, 40, chk_step=16)
    H.close()
    assert rc == _lib.TBG_ENORM and fs == 16
    # 524,288 orbitals with 16 complex64 states need 256 MiB: refused
    H, bes, psi = _case(512, 16, 4, "fp32", 14)
    rc, _, _ = _noalloc(H, bes, psi, 4)
    H.close()
    assert rc == _lib.TBG_EINVAL
    assert b"too large" in _noalloc.msg
