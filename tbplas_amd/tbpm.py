"""Kernel-level Python API over the C ABI: prepared Hamiltonians and the
correlation runs, with device tensors as storage.

These functions take and return numpy arrays (host) for convenience; the
arrays are staged through device tensors and the library works on those.
"""
import ctypes
import math

import numpy as np

from . import _lib

__all__ = ["Hamiltonian", "DosRun", "bessel_series", "fermi_cheb_coef", "corr_dos",
           "corr_dos_batch", "corr_ac", "corr_dc", "corr_ldos", "ldos_states", "psi_t",
           "quasi_eigenstates", "corr_dyn_pol", "hall_mu", "haydock_coef", "haydock_green",
           "cheb_step", "cheb_poly",
           "precision_code"]


def precision_code(precision):
    if precision in ("fp64", _lib.TBG_FP64):
        return _lib.TBG_FP64
    if precision in ("fp32", _lib.TBG_FP32):
        return _lib.TBG_FP32
    raise ValueError(f"unknown precision {precision!r}")


class Hamiltonian:
    """A prepared device Hamiltonian (tbg_ham_create).

    indptr/indices/hop may be numpy arrays or device tensors (int64, int64,
    complex128); dx/dy (float64) enable the current operators
    J = (0, cur_scale) * hop * d (funcs.f90:12-33).  ``order`` (optional,
    int64 (n,)) lists the rows in the order the device layout stores them,
    e.g. ``Sample.tile_order()``; ``tiles`` (optional, int64 row offsets
    0..n in that order) marks spatial tiles, which enables three Chebyshev
    terms per pass.  Results are in the caller's row order either way.
    """

    def __init__(self, n, indptr, indices, hop, dx=None, dy=None, cur_scale=1.0,
                 precision="fp64", handle=None, device=0, order=None, tiles=None):
        torch = _lib.torch_cuda()
        self.handle = handle or _lib.default_handle(device)
        dv = f"cuda:{self.handle.device}"

        def to_dev(a, dt):
            if a is None:
                return None
            if isinstance(a, torch.Tensor):
                return a.to(device=dv, dtype=dt).contiguous()
            return torch.from_numpy(np.ascontiguousarray(a)).to(device=dv, dtype=dt)

        ip = to_dev(indptr, torch.int64)
        ix = to_dev(indices, torch.int64)
        hv = to_dev(hop, torch.complex128)
        dxd = to_dev(dx, torch.float64)
        dyd = to_dev(dy, torch.float64)
        if ix.numel() == 0:
            ix = torch.zeros(1, dtype=torch.int64, device=dv)
            hv = torch.zeros(1, dtype=torch.complex128, device=dv)
        self.n = int(n)
        self.precision = precision_code(precision)
        self.has_current = dx is not None
        od = to_dev(order, torch.int64)
        if od is not None and od.numel() != self.n:
            raise ValueError("order must list every row once")
        self.ordered = od is not None
        tp = None if tiles is None else np.ascontiguousarray(tiles, dtype=np.int64)
        self.tiled = tp is not None
        self.ntiles = 0 if tp is None else len(tp) - 1
        out = ctypes.c_void_p()
        _lib.check(_lib.lib().tbg_ham_create_ordered(
            self.handle.ptr, self.n, _lib.ptr(ip), _lib.ptr(ix), _lib.ptr(hv),
            _lib.ptr(dxd) if dxd is not None else None,
            _lib.ptr(dyd) if dyd is not None else None,
            float(cur_scale), self.precision,
            _lib.ptr(od) if od is not None else None,
            tp.ctypes.data if tp is not None else None,
            0 if tp is None else len(tp) - 1, ctypes.byref(out)))
        self._h = out

    @property
    def ptr(self):
        return self._h

    def info(self):
        n, nnz, pad = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
        w = ctypes.c_int()
        _lib.check(_lib.lib().tbg_ham_info(self._h, ctypes.byref(n), ctypes.byref(nnz),
                                           ctypes.byref(pad), ctypes.byref(w)))
        real = ctypes.c_int()
        _lib.check(_lib.lib().tbg_ham_is_real(self._h, ctypes.byref(real)))
        return {"n": n.value, "nnz": nnz.value, "padded": pad.value,
                "max_width": w.value, "real_values": bool(real.value)}

    def blocking_info(self, states, precision=None):
        """(workgroup tiles, halo factor) of the three-term pass for
        ``states`` vectors; (0, 0.0) when H runs one term per pass."""
        nb, halo = ctypes.c_int(), ctypes.c_double()
        prec = self.precision if precision is None else precision_code(precision)
        _lib.check(_lib.lib().tbg_ham_blocking_info(self.handle.ptr, self._h, int(states), prec,
                                                    ctypes.byref(nb), ctypes.byref(halo)))
        return nb.value, halo.value

    def blocking_launch(self, states, precision=None):
        """(rows of the widest workgroup tile, threads per workgroup) of the
        three-term pass for ``states`` vectors; (0, 0) when H runs one term
        per pass."""
        rows, nt = ctypes.c_int(), ctypes.c_int()
        prec = self.precision if precision is None else precision_code(precision)
        _lib.check(_lib.lib().tbg_ham_blocking_launch(self.handle.ptr, self._h, int(states),
                                                      prec, ctypes.byref(rows),
                                                      ctypes.byref(nt)))
        return rows.value, nt.value

    def blocking_bytes(self, states, precision=None):
        """What one full tile pass must move for ``states`` vectors:
        {"tables": bytes, "gathered_rows", "streamed_rows", "owned_rows"} and
        "row_bytes" (slots x element bytes) and "pass_bytes" = tables +
        row_bytes x (gathered + streamed + 4 owned); None when H runs one
        term per pass."""
        prec = self.precision if precision is None else precision_code(precision)
        parts = np.zeros(4, dtype=np.int64)
        _lib.check(_lib.lib().tbg_ham_blocking_bytes(self.handle.ptr, self._h, int(states), prec,
                                                     parts.ctypes.data))
        if parts[3] == 0:
            return None
        slots = 1
        while slots < states:
            slots <<= 1
        p3 = 3
        while p3 < states:
            p3 <<= 1
        slots = min(slots, p3)
        rb = slots * (8 if prec == _lib.TBG_FP32 else 16)
        tab, gat, stream, own = (int(x) for x in parts)
        return {"tables": tab, "gathered_rows": gat, "streamed_rows": stream,
                "owned_rows": own, "row_bytes": rb,
                "pass_bytes": tab + rb * (gat + stream + 4 * own)}

    def close(self):
        if getattr(self, "_h", None):
            _lib.lib().tbg_ham_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def bessel_series(time_scaled, precision=1e-14, bessel_max=250):
    """Solver._get_bessel_series (solver.py:194-224), scipy.special.jv."""
    import scipy.special as spec
    series = []
    for i in range(bessel_max):
        v = spec.jv(i, time_scaled)
        if abs(v) > precision:
            series.append(float(v))
        elif abs(spec.jv(i + 1, time_scaled)) > precision:
            series.append(float(v))
        else:
            return series
    raise ValueError("Bessel_max too low")


def fermi_cheb_coef(nr_fermi, beta, mu, one_minus, eps):
    """Chebyshev coefficients of the (one-minus) Fermi operator
    (get_Fermi_cheb_coef, funcs.f90:60-123): FFT of f(cos(2 pi i / N)),
    Jackson damping (kpm.f90:8-29), cut at the first two consecutive
    coefficients below eps."""
    nr = int(nr_fermi)
    if nr < 4 or nr & (nr - 1):
        raise ValueError("nr_Fermi_fft_steps must be a power of two >= 4")
    r0 = 2 * math.pi / nr
    energy = np.cos(np.arange(nr) * r0)
    f = np.empty(nr)
    hi = energy >= mu
    x = np.exp(beta * (mu - energy[hi]))
    f[hi] = x / (1 + x)
    x = np.exp(beta * (energy[~hi] - mu))
    f[~hi] = 1 / (1 + x)
    if one_minus:
        f = 1.0 - f
    c = np.fft.fft(f.astype(np.complex128))
    n = nr // 2
    k = np.arange(1, n)
    q = math.pi / n
    cotq = math.cos(q) / math.sin(q)
    g = np.empty(n)
    g[0] = 0.5
    g[1:] = ((n - k) * np.cos(q * k) + cotq * np.sin(q * k)) / n
    coef = g * c.real[:n] / n
    small = np.abs(coef) < eps
    nxt = np.append(small[1:], True)
    hit = np.nonzero(small & nxt)[0]
    keep = int(hit[0]) + 1 if len(hit) else n
    return coef[:keep].copy()


def _states_dev(psi0, device):
    psi0 = np.ascontiguousarray(np.atleast_2d(psi0), dtype=np.complex128)
    return _lib.dev(psi0, np.complex128, device), psi0.shape


def corr_dos(H, bes, psi0, n_timestep, chk_step=128, chk_thr=1e-9):
    """Per-state DOS correlation (K, T+1) (tbpm_dos, tbpm.f90:6-93)."""
    torch = _lib.torch_cuda()
    dv = H.handle.device
    p, (K, n) = _states_dev(psi0, dv)
    if n != H.n:
        raise ValueError("state length does not match the Hamiltonian")
    bes = np.ascontiguousarray(bes, dtype=np.float64)
    out = torch.empty((K, n_timestep + 1), dtype=torch.complex128, device=f"cuda:{dv}")
    fs = ctypes.c_int(0)
    rc = _lib.lib().tbg_corr_dos(H.handle.ptr, H.ptr, bes.ctypes.data, len(bes), _lib.ptr(p),
                                 K, int(n_timestep), int(chk_step), float(chk_thr),
                                 _lib.ptr(out), ctypes.byref(fs))
    _lib.check(rc, fs)
    return out.cpu().numpy()


def ldos_states(psi, site_indices):
    """LDOS initial states: each random state kept on the selected sites and
    normalised (tbpm.f90:151-156).  psi (K, n) -> (K, n) complex128."""
    psi = np.atleast_2d(np.asarray(psi, dtype=np.complex128))
    out = np.zeros_like(psi)
    idx = np.asarray(site_indices, dtype=np.int64)
    out[:, idx] = psi[:, idx]
    norm = np.sqrt(np.einsum("kj,kj->k", out.conj(), out).real)
    return out / norm[:, None]


def corr_ldos(H, bes, psi0, n_timestep, chk_step=128, chk_thr=1e-9):
    """Per-state LDOS correlation (K, T+1) <psi0|psi(t)> (tbpm_ldos,
    tbpm.f90:97-189); psi0 are LDOS states (``ldos_states``)."""
    torch = _lib.torch_cuda()
    dv = H.handle.device
    p, (K, n) = _states_dev(psi0, dv)
    if n != H.n:
        raise ValueError("state length does not match the Hamiltonian")
    bes = np.ascontiguousarray(bes, dtype=np.float64)
    out = torch.empty((K, n_timestep + 1), dtype=torch.complex128, device=f"cuda:{dv}")
    fs = ctypes.c_int(0)
    rc = _lib.lib().tbg_corr_ldos(H.handle.ptr, H.ptr, bes.ctypes.data, len(bes), _lib.ptr(p),
                                  K, int(n_timestep), int(chk_step), float(chk_thr),
                                  _lib.ptr(out), ctypes.byref(fs))
    _lib.check(rc, fs)
    return out.cpu().numpy()


def psi_t(H, bes, psi0, time_log, n_timestep, chk_step=128, chk_thr=1e-9):
    """Wave function psi0 (n,) propagated n_timestep steps, logged after the
    steps in time_log (strictly increasing, in [0, T]): (n_log, n)
    complex128 (tbpm_psi_t, tbpm.f90:1029-1093)."""
    torch = _lib.torch_cuda()
    dv = H.handle.device
    p, (K, n) = _states_dev(np.atleast_2d(psi0), dv)
    if K != 1 or n != H.n:
        raise ValueError("psi0 must be one state of the Hamiltonian's size")
    tl = np.ascontiguousarray(time_log, dtype=np.int64)
    bes = np.ascontiguousarray(bes, dtype=np.float64)
    out = torch.zeros((len(tl), n), dtype=torch.complex128, device=f"cuda:{dv}")
    fs = ctypes.c_int(0)
    rc = _lib.lib().tbg_psi_t(H.handle.ptr, H.ptr, bes.ctypes.data, len(bes), _lib.ptr(p),
                              tl.ctypes.data, len(tl), int(n_timestep), int(chk_step),
                              float(chk_thr), _lib.ptr(out), ctypes.byref(fs))
    _lib.check(rc, fs)
    return out.cpu().numpy()


def corr_dyn_pol(H, bes, coef_F, coef_omF, psi0, n_timestep, q_points, orb_pos,
                 chk_step=128, chk_thr=1e-9):
    """Per-state dynamic-polarisation correlation (K, n_q, T) float64
    (tbpm_dyn_pol, tbpm.f90:355-493); orb_pos (n, 3) in the caller's order,
    q_points (n_q, 3) in the inverse units of orb_pos."""
    torch = _lib.torch_cuda()
    dv = H.handle.device
    p, (K, n) = _states_dev(psi0, dv)
    if n != H.n:
        raise ValueError("state length does not match the Hamiltonian")
    bes = np.ascontiguousarray(bes, dtype=np.float64)
    cF = np.ascontiguousarray(coef_F, dtype=np.float64)
    cO = np.ascontiguousarray(coef_omF, dtype=np.float64)
    q = np.ascontiguousarray(np.atleast_2d(q_points), dtype=np.float64)
    pos = np.ascontiguousarray(orb_pos, dtype=np.float64)
    if q.shape[1] != 3 or pos.shape != (n, 3):
        raise ValueError("q_points must be (n_q, 3) and orb_pos (n, 3)")
    out = torch.empty((K, len(q), n_timestep), dtype=torch.float64, device=f"cuda:{dv}")
    fs = ctypes.c_int(0)
    rc = _lib.lib().tbg_corr_dyn_pol(H.handle.ptr, H.ptr, bes.ctypes.data, len(bes),
                                     cF.ctypes.data, len(cF), cO.ctypes.data, len(cO),
                                     _lib.ptr(p), K, int(n_timestep), q.ctypes.data, len(q),
                                     pos.ctypes.data, int(chk_step), float(chk_thr),
                                     _lib.ptr(out), ctypes.byref(fs))
    _lib.check(rc, fs)
    return out.cpu().numpy()


def haydock_coef(H, psi0, depth, rescale):
    """Haydock recursion coefficients (Haydock_coef, propagation.f90:85-155)
    of each starting state psi0 (K, n) under rescale * H~: (a (K, depth)
    complex128, b (K, depth) float64).  H: complex128."""
    dv = H.handle.device
    p, (K, n) = _states_dev(psi0, dv)
    if n != H.n:
        raise ValueError("state length does not match the Hamiltonian")
    a = np.zeros((K, int(depth)), dtype=np.complex128)
    b = np.zeros((K, int(depth)), dtype=np.float64)
    _lib.check(_lib.lib().tbg_haydock_coef(H.handle.ptr, H.ptr, _lib.ptr(p), K, int(depth),
                                           float(rescale), a.ctypes.data, b.ctypes.data))
    return a, b


def haydock_green(coefa, coefb, energies, delta):
    """G00(E + i delta) by the continued fraction of green_function
    (funcs.f90:151-172), for all energies at once."""
    z = np.asarray(energies, dtype=np.float64) + 1j * delta
    g = np.zeros_like(z)
    for i in range(len(coefa) - 1, -1, -1):
        g = 1.0 / (z - coefa[i] - coefb[i] ** 2 * g)
    return g


def hall_mu(H, psi0, n_kernel, direction=1):
    """Per-state Kubo-Bastin moments g_i g_j <T_i(H) J_x psi | J_b T_j(H) psi>
    (tbpm_kbdc, tbpm.f90:836-960): (K, n_kernel, n_kernel) complex128,
    [s, i, j]; direction 1: b = x, 2: b = y.  H: complex128 with currents."""
    torch = _lib.torch_cuda()
    dv = H.handle.device
    p, (K, n) = _states_dev(psi0, dv)
    if n != H.n:
        raise ValueError("state length does not match the Hamiltonian")
    out = torch.empty((K, n_kernel, n_kernel), dtype=torch.complex128, device=f"cuda:{dv}")
    _lib.check(_lib.lib().tbg_hall_mu(H.handle.ptr, H.ptr, _lib.ptr(p), K, int(n_kernel),
                                      int(direction), _lib.ptr(out)))
    # each state is column-major (i + j N): the array read row-major is [s, j, i]
    return out.cpu().numpy().transpose(0, 2, 1).copy()


def quasi_eigenstates(H, bes, psi0, n_timestep, t_step, energies, chk_step=128, chk_thr=1e-9):
    """Per-state |QE|^2 (K, n_en, n) of the quasi-eigenstates at the given
    energies (tbpm_eigenstates, tbpm.f90:721-830); the reference's output
    is their average over states."""
    torch = _lib.torch_cuda()
    dv = H.handle.device
    p, (K, n) = _states_dev(psi0, dv)
    if n != H.n:
        raise ValueError("state length does not match the Hamiltonian")
    bes = np.ascontiguousarray(bes, dtype=np.float64)
    en = np.ascontiguousarray(energies, dtype=np.float64)
    out = torch.empty((K, len(en), n), dtype=torch.float64, device=f"cuda:{dv}")
    fs = ctypes.c_int(0)
    rc = _lib.lib().tbg_quasi_eigenstates(H.handle.ptr, H.ptr, bes.ctypes.data, len(bes),
                                          _lib.ptr(p), K, int(n_timestep), float(t_step),
                                          en.ctypes.data, len(en), int(chk_step),
                                          float(chk_thr), _lib.ptr(out), ctypes.byref(fs))
    _lib.check(rc, fs)
    return out.cpu().numpy()


def corr_dos_batch(hams, bes, psi0_list, n_timestep, chk_step=128, chk_thr=1e-9):
    """DOS correlations of several Hamiltonians in one run per term."""
    torch = _lib.torch_cuda()
    handle = hams[0].handle
    dv = handle.device
    bes = np.ascontiguousarray(bes, dtype=np.float64)
    ps, outs = [], []
    K = None
    for H, psi in zip(hams, psi0_list):
        p, (k, n) = _states_dev(psi, dv)
        if n != H.n:
            raise ValueError("state length does not match the Hamiltonian")
        K = k if K is None else K
        if k != K:
            raise ValueError("every instance needs the same number of states")
        ps.append(p)
        outs.append(torch.empty((K, n_timestep + 1), dtype=torch.complex128, device=f"cuda:{dv}"))
    ni = len(hams)
    harr = (ctypes.c_void_p * ni)(*[H.ptr.value for H in hams])
    parr = (ctypes.c_void_p * ni)(*[t.data_ptr() for t in ps])
    oarr = (ctypes.c_void_p * ni)(*[t.data_ptr() for t in outs])
    fi, fs = ctypes.c_int(-1), ctypes.c_int(0)
    rc = _lib.lib().tbg_corr_dos_batch(handle.ptr, ni, harr, bes.ctypes.data, len(bes), parr, K,
                                       int(n_timestep), int(chk_step), float(chk_thr), oarr,
                                       ctypes.byref(fi), ctypes.byref(fs))
    _lib.check(rc, fs)
    return [o.cpu().numpy() for o in outs]


class DosRun:
    """A DOS correlation run advanced in pieces (tbg_dos_begin/advance/end).

    ``advance`` returns the device time of the steps in milliseconds,
    measured with HIP events on the library's stream.
    """

    def __init__(self, hams, bes, psi0_list, n_timestep):
        torch = _lib.torch_cuda()
        if isinstance(hams, Hamiltonian):
            hams, psi0_list = [hams], [psi0_list]
        self.hams = list(hams)
        self.handle = self.hams[0].handle
        dv = self.handle.device
        self.bes = np.ascontiguousarray(bes, dtype=np.float64)
        self.ps = []
        for H, psi in zip(self.hams, psi0_list):
            p, (k, n) = _states_dev(psi, dv)
            if n != H.n:
                raise ValueError("state length does not match the Hamiltonian")
            self.ps.append(p)
        self.K = self.ps[0].shape[0]
        self.T = int(n_timestep)
        self.done = 0
        ni = len(self.hams)
        harr = (ctypes.c_void_p * ni)(*[H.ptr.value for H in self.hams])
        parr = (ctypes.c_void_p * ni)(*[t.data_ptr() for t in self.ps])
        out = ctypes.c_void_p()
        _lib.check(_lib.lib().tbg_dos_begin(self.handle.ptr, ni, harr, self.bes.ctypes.data,
                                            len(self.bes), parr, self.K, self.T,
                                            ctypes.byref(out)))
        self._r = out
        self._torch = torch

    def advance(self, nsteps):
        ms = ctypes.c_double(0.0)
        _lib.check(_lib.lib().tbg_run_advance(self._r, int(nsteps), ctypes.byref(ms)))
        self.done += int(nsteps)
        return ms.value

    def blocking_bytes(self):
        """Hamiltonian.blocking_bytes for the tables this run uses (a
        batch's concatenated tiles included); None without tile passes."""
        parts = np.zeros(4, dtype=np.int64)
        rb = ctypes.c_int(0)
        _lib.check(_lib.lib().tbg_run_blocking_bytes(self._r, parts.ctypes.data,
                                                     ctypes.byref(rb)))
        if parts[3] == 0:
            return None
        tab, gat, stream, own = (int(x) for x in parts)
        return {"tables": tab, "gathered_rows": gat, "streamed_rows": stream,
                "owned_rows": own, "row_bytes": rb.value,
                "pass_bytes": tab + rb.value * (gat + stream + 4 * own)}

    def result(self, chk_step=128, chk_thr=1e-9):
        torch = self._torch
        dv = self.handle.device
        outs = [torch.zeros((self.K, self.T + 1), dtype=torch.complex128, device=f"cuda:{dv}")
                for _ in self.hams]
        oarr = (ctypes.c_void_p * len(outs))(*[o.data_ptr() for o in outs])
        fi, fs = ctypes.c_int(-1), ctypes.c_int(0)
        rc = _lib.lib().tbg_dos_end(self._r, int(chk_step), float(chk_thr), oarr,
                                    ctypes.byref(fi), ctypes.byref(fs))
        _lib.check(rc, fs)
        res = [o.cpu().numpy() for o in outs]
        return res if len(res) > 1 else res[0]

    def close(self):
        if getattr(self, "_r", None):
            _lib.lib().tbg_run_destroy(self._r)
            self._r = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def corr_ac(H, bes, coef_F, coef_omF, psi0, n_timestep, chk_step=128, chk_thr=1e-9):
    """Per-state AC correlation (K, 4, T) (tbpm_accond, tbpm.f90:193-351)."""
    torch = _lib.torch_cuda()
    dv = H.handle.device
    p, (K, n) = _states_dev(psi0, dv)
    bes = np.ascontiguousarray(bes, dtype=np.float64)
    cF = np.ascontiguousarray(coef_F, dtype=np.float64)
    cO = np.ascontiguousarray(coef_omF, dtype=np.float64)
    out = torch.empty((K, 4, n_timestep), dtype=torch.complex128, device=f"cuda:{dv}")
    fs = ctypes.c_int(0)
    rc = _lib.lib().tbg_corr_ac(H.handle.ptr, H.ptr, bes.ctypes.data, len(bes), cF.ctypes.data,
                                len(cF), cO.ctypes.data, len(cO), _lib.ptr(p), K,
                                int(n_timestep), int(chk_step), float(chk_thr), _lib.ptr(out),
                                ctypes.byref(fs))
    _lib.check(rc, fs)
    return out.cpu().numpy()


def corr_dc(H, bes, psi0, n_timestep, t_step, energies, en_inds, chk_step=128,
            chk_thr=1e-9):
    """Per-state DC correlations (tbpm_dccond, tbpm.f90:497-718):
    dos (K, T) and dc (K, 2, n_en, T)."""
    torch = _lib.torch_cuda()
    dv = H.handle.device
    p, (K, n) = _states_dev(psi0, dv)
    bes = np.ascontiguousarray(bes, dtype=np.float64)
    energies = np.ascontiguousarray(energies, dtype=np.float64)
    en_inds = np.ascontiguousarray(en_inds, dtype=np.int64)
    n_en = len(en_inds)
    dos = torch.empty((K, n_timestep), dtype=torch.complex128, device=f"cuda:{dv}")
    dc = torch.empty((K, 2, max(n_en, 1), n_timestep), dtype=torch.complex128,
                     device=f"cuda:{dv}")
    fs = ctypes.c_int(0)
    rc = _lib.lib().tbg_corr_dc(H.handle.ptr, H.ptr, bes.ctypes.data, len(bes), _lib.ptr(p), K,
                                int(n_timestep), float(t_step), energies.ctypes.data,
                                en_inds.ctypes.data, n_en, int(chk_step), float(chk_thr),
                                _lib.ptr(dos), _lib.ptr(dc), ctypes.byref(fs))
    _lib.check(rc, fs)
    return dos.cpu().numpy(), dc.cpu().numpy()[:, :, :n_en, :]


def cheb_step(H, bes, psi, fwd=True):
    """psi (K, n) -> exp(-+i H~ tau) psi (cheb_wf_timestep)."""
    dv = H.handle.device
    p, (K, n) = _states_dev(psi, dv)
    bes = np.ascontiguousarray(bes, dtype=np.float64)
    _lib.check(_lib.lib().tbg_cheb_step(H.handle.ptr, H.ptr, bes.ctypes.data, len(bes),
                                        _lib.ptr(p), K, 1 if fwd else 0))
    return p.cpu().numpy()


def cheb_poly(H, coef, psi):
    """psi (K, n) -> sum_k c_k T_k(H~) psi (Fermi, propagation.f90:52-82)."""
    dv = H.handle.device
    p, (K, n) = _states_dev(psi, dv)
    coef = np.ascontiguousarray(coef, dtype=np.float64)
    _lib.check(_lib.lib().tbg_cheb_poly(H.handle.ptr, H.ptr, coef.ctypes.data, len(coef),
                                        _lib.ptr(p), K))
    return p.cpu().numpy()
