"""Oracle implementation -- test infrastructure only (see package docstring)."""
import ctypes
import math
import os
import subprocess

import numpy as np

__all__ = [
    "oracle_lib", "ref_lib", "ref_available",
    "orc_tbpm_dos", "orc_tbpm_accond", "orc_tbpm_dccond", "orc_step",
    "orc_fermi_coef", "orc_csr_mv", "orc_tbpm_ldos", "orc_psi_t", "orc_quasi_eigenstates",
    "orc_cheb_apply", "orc_tbpm_dyn_pol", "ref_tbpm_dyn_pol",
    "jackson_kernel", "orc_tbpm_kbdc", "ref_tbpm_kbdc", "orc_cond_from_trace",
    "ref_cond_from_trace", "orc_haydock_coef", "orc_green_function", "ref_ldos_haydock",
    "orc_supercell_hops",
    "ref_random_state", "ref_tbpm_dos", "ref_tbpm_accond", "ref_tbpm_dccond",
    "ref_fermi_coef", "ref_tbpm_ldos", "ref_tbpm_psi_t", "ref_tbpm_eigenstates",
    "bessel_series", "build_ham_csr", "build_ham_dxy_loop",
    "build_ham_dxy_stable", "graphene_hops_loop", "graphene_hops_vec", "get_rescale_loop",
    "set_mag_field_loop", "kspace_graphene_corr", "kspace_graphene_corr_multi", "dense_corr",
    "kspace_graphene_ac", "cheb_series_values",
    "analyzer_calc_dos", "analyzer_calc_ac_cond", "analyzer_calc_dc_cond",
    "analyzer_calc_diff_coeff",
    "window_hanning", "window_exp",
]

_HERE = os.path.dirname(os.path.abspath(__file__))
_ORC_SO = os.path.join(_HERE, "_build", "liboracle.so")
_REF_SO = os.path.join(_HERE, "_ref", "libtbpm_ref.so")
_c_i64p = ctypes.POINTER(ctypes.c_int64)
_c_dp = ctypes.POINTER(ctypes.c_double)
_c_ip = ctypes.POINTER(ctypes.c_int)
_libs = {}


def _ptr(a, ctype=ctypes.c_double):
    return a.ctypes.data_as(ctypes.POINTER(ctype))


def oracle_lib():
    """Load (building if needed) the C restatement library."""
    if "orc" not in _libs:
        if not os.path.exists(_ORC_SO):
            subprocess.run(["make", "-C", _HERE, "oracle"], check=True,
                           stdout=subprocess.DEVNULL)
        _libs["orc"] = ctypes.CDLL(_ORC_SO)
    return _libs["orc"]


def ref_available():
    return os.path.exists(_REF_SO)


def ref_lib():
    """Load the reference Fortran TBPM library compiled from its sources."""
    if "ref" not in _libs:
        if not ref_available():
            raise FileNotFoundError(
                f"{_REF_SO} missing: run `make -C oracle ref` where the "
                "reference sources exist")
        _libs["ref"] = ctypes.CDLL(_REF_SO)
    return _libs["ref"]


def _csr_args(indptr, indices, hop):
    indptr = np.ascontiguousarray(indptr, dtype=np.int64)
    indices = np.ascontiguousarray(indices, dtype=np.int64)
    hop = np.ascontiguousarray(hop, dtype=np.complex128)
    return indptr, indices, hop


def _check_rc(rc, fail_step, what):
    if rc == 1:
        raise FloatingPointError(f"{what}: wave function norm broken at step "
                                 f"{fail_step.value}")
    if rc != 0:
        raise RuntimeError(f"{what}: oracle returned {rc}")


# ---------------------------------------------------------------------------
# C restatement (oracle/tbpm_oracle.c)
# ---------------------------------------------------------------------------
def orc_tbpm_dos(indptr, indices, hop, bes, psi0, n_timestep,
                 chk_step=128, chk_thr=1e-9):
    """tbpm_dos (tbpm.f90:6-93) with injected psi0 (S, n).

    Returns the per-sample correlation (S, T+1); the reference's output is
    the running average ``sum_s c_s / S`` of these rows.
    """
    lib = oracle_lib()
    indptr, indices, hop = _csr_args(indptr, indices, hop)
    psi0 = np.ascontiguousarray(np.atleast_2d(psi0), dtype=np.complex128)
    bes = np.ascontiguousarray(bes, dtype=np.float64)
    S, n = psi0.shape
    corr = np.zeros((S, n_timestep + 1), dtype=np.complex128)
    fs = ctypes.c_int(0)
    rc = lib.orc_tbpm_dos(
        ctypes.c_int64(n), _ptr(indptr, ctypes.c_int64),
        _ptr(indices, ctypes.c_int64), _ptr(hop), _ptr(bes),
        ctypes.c_int(len(bes)), _ptr(psi0), ctypes.c_int(S),
        ctypes.c_int(n_timestep), ctypes.c_int(chk_step),
        ctypes.c_double(chk_thr), _ptr(corr), ctypes.byref(fs))
    _check_rc(rc, fs, "orc_tbpm_dos")
    return corr


def orc_tbpm_accond(indptr, indices, hop, dx, dy, rescale, beta_re, mu_re,
                    bes, psi0, n_timestep, nr_fermi=2 ** 15,
                    fermi_prec=1e-12, chk_step=128, chk_thr=1e-9):
    """tbpm_accond (tbpm.f90:193-351); per-sample (S, 4, T)."""
    lib = oracle_lib()
    indptr, indices, hop = _csr_args(indptr, indices, hop)
    dx = np.ascontiguousarray(dx, dtype=np.float64)
    dy = np.ascontiguousarray(dy, dtype=np.float64)
    psi0 = np.ascontiguousarray(np.atleast_2d(psi0), dtype=np.complex128)
    bes = np.ascontiguousarray(bes, dtype=np.float64)
    S, n = psi0.shape
    corr = np.zeros((S, 4, n_timestep), dtype=np.complex128)
    fs = ctypes.c_int(0)
    rc = lib.orc_tbpm_accond(
        ctypes.c_int64(n), _ptr(indptr, ctypes.c_int64),
        _ptr(indices, ctypes.c_int64), _ptr(hop), _ptr(dx), _ptr(dy),
        ctypes.c_double(rescale), ctypes.c_double(beta_re),
        ctypes.c_double(mu_re), ctypes.c_int64(nr_fermi),
        ctypes.c_double(fermi_prec), _ptr(bes), ctypes.c_int(len(bes)),
        _ptr(psi0), ctypes.c_int(S), ctypes.c_int(n_timestep),
        ctypes.c_int(chk_step), ctypes.c_double(chk_thr), _ptr(corr),
        ctypes.byref(fs))
    _check_rc(rc, fs, "orc_tbpm_accond")
    return corr


def orc_tbpm_dccond(indptr, indices, hop, dx, dy, rescale, bes, psi0,
                    n_timestep, t_step, energies, en_inds, chk_step=128,
                    chk_thr=1e-9):
    """tbpm_dccond (tbpm.f90:497-718).

    Returns per-sample ``dos (S, T)`` and ``dc (S, 2, n_en, T)``.
    """
    lib = oracle_lib()
    indptr, indices, hop = _csr_args(indptr, indices, hop)
    dx = np.ascontiguousarray(dx, dtype=np.float64)
    dy = np.ascontiguousarray(dy, dtype=np.float64)
    psi0 = np.ascontiguousarray(np.atleast_2d(psi0), dtype=np.complex128)
    bes = np.ascontiguousarray(bes, dtype=np.float64)
    energies = np.ascontiguousarray(energies, dtype=np.float64)
    en_inds = np.ascontiguousarray(en_inds, dtype=np.int64)
    S, n = psi0.shape
    n_en = len(en_inds)
    dos = np.zeros((S, n_timestep), dtype=np.complex128)
    dc = np.zeros((S, 2, n_en, n_timestep), dtype=np.complex128)
    fs = ctypes.c_int(0)
    rc = lib.orc_tbpm_dccond(
        ctypes.c_int64(n), _ptr(indptr, ctypes.c_int64),
        _ptr(indices, ctypes.c_int64), _ptr(hop), _ptr(dx), _ptr(dy),
        ctypes.c_double(rescale), _ptr(bes), ctypes.c_int(len(bes)),
        _ptr(psi0), ctypes.c_int(S), ctypes.c_int(n_timestep),
        ctypes.c_double(t_step), _ptr(energies),
        _ptr(en_inds, ctypes.c_int64), ctypes.c_int(n_en),
        ctypes.c_int(chk_step), ctypes.c_double(chk_thr), _ptr(dos),
        _ptr(dc), ctypes.byref(fs))
    _check_rc(rc, fs, "orc_tbpm_dccond")
    return dos, dc


def orc_step(indptr, indices, hop, bes, psi, fwd=True):
    """One cheb_wf_timestep (propagation.f90:12-49)."""
    lib = oracle_lib()
    indptr, indices, hop = _csr_args(indptr, indices, hop)
    psi = np.ascontiguousarray(psi, dtype=np.complex128)
    bes = np.ascontiguousarray(bes, dtype=np.float64)
    out = np.zeros_like(psi)
    rc = lib.orc_step(ctypes.c_int64(len(psi)), _ptr(indptr, ctypes.c_int64),
                      _ptr(indices, ctypes.c_int64), _ptr(hop), _ptr(bes),
                      ctypes.c_int(len(bes)), _ptr(psi), _ptr(out),
                      ctypes.c_int(1 if fwd else 0))
    if rc:
        raise RuntimeError(f"orc_step returned {rc}")
    return out


def orc_tbpm_ldos(indptr, indices, hop, bes, psi0, n_timestep):
    """tbpm_ldos (tbpm.f90:97-189) on caller-supplied LDOS states psi0
    (S, n) (already kept on the sites and normalised, tbpm.f90:151-156):
    per-state <psi0|psi(t)>, (S, T+1).  Built from orc_step."""
    psi0 = np.atleast_2d(np.asarray(psi0, dtype=np.complex128))
    corr = np.zeros((psi0.shape[0], n_timestep + 1), dtype=np.complex128)
    for s, w0 in enumerate(psi0):
        wt = w0.copy()
        corr[s, 0] = np.vdot(w0, wt)
        for k in range(1, n_timestep + 1):
            wt = orc_step(indptr, indices, hop, bes, wt, True)
            corr[s, k] = np.vdot(w0, wt)
    return corr


def orc_psi_t(indptr, indices, hop, bes, psi0, time_log, n_timestep):
    """tbpm_psi_t (tbpm.f90:1029-1093): psi0 propagated, logged after the
    steps in time_log (sorted; the log pointer stays on the last entry):
    (n_log, n).  Built from orc_step."""
    tl = [int(t) for t in time_log]
    out = np.zeros((len(tl), len(psi0)), dtype=np.complex128)
    wt = np.asarray(psi0, dtype=np.complex128).copy()
    ptr = 0
    for t in range(1, n_timestep + 1):
        if tl[ptr] == t - 1:
            out[ptr] = wt
            if ptr < len(tl) - 1:
                ptr += 1
        wt = orc_step(indptr, indices, hop, bes, wt, True)
    if tl[ptr] == n_timestep:
        out[ptr] = wt
    return out


def orc_quasi_eigenstates(indptr, indices, hop, bes, psi0, n_timestep, t_step, energies):
    """tbpm_eigenstates (tbpm.f90:721-830) on caller-supplied states psi0
    (S, n): QE_i = psi0 + sum_k W(k) (e^{i E_i k dt} psi(+k) + e^{-i E_i k
    dt} psi(-k)), W = Hanning 0.5 (1 + cos(pi k / T)), normalised; returns
    the per-state |QE|^2, (S, n_E, n) (the reference averages them)."""
    psi0 = np.atleast_2d(np.asarray(psi0, dtype=np.complex128))
    en = np.asarray(energies, dtype=np.float64)
    out = np.zeros((psi0.shape[0], len(en), psi0.shape[1]))
    for s, w0 in enumerate(psi0):
        pos, neg = w0.copy(), w0.copy()
        wfq = np.tile(w0, (len(en), 1))
        for k in range(1, n_timestep + 1):
            pos = orc_step(indptr, indices, hop, bes, pos, True)
            neg = orc_step(indptr, indices, hop, bes, neg, False)
            W = 0.5 * (1 + math.cos(math.pi * k / n_timestep))
            for i, e in enumerate(en):
                P = complex(math.cos(e * k * t_step), math.sin(e * k * t_step))
                wfq[i] = wfq[i] + P * pos * W
                wfq[i] = wfq[i] + P.conjugate() * neg * W
        for i in range(len(en)):
            wfq[i] = wfq[i] / np.sqrt(np.vdot(wfq[i], wfq[i]).real)
        out[s] = np.abs(wfq) ** 2
    return out


def orc_cheb_apply(indptr, indices, hop, coef, psi):
    """Fermi (propagation.f90:52-82): sum_k c_k T_k(H) psi by the recurrence
    T_0 = psi, T_1 = H psi, T_k = 2 H T_{k-1} - T_{k-2}, in that order of
    accumulation; H products from orc_csr_mv."""
    p0 = np.asarray(psi, dtype=np.complex128).copy()
    out = coef[0] * p0
    if len(coef) == 1:
        return out
    p1 = orc_csr_mv(indptr, indices, hop, p0)
    out = out + coef[1] * p1
    for k in range(2, len(coef)):
        p2 = 2.0 * orc_csr_mv(indptr, indices, hop, p1) - p0
        out = out + coef[k] * p2
        p0, p1 = p1, p2
    return out


def orc_tbpm_dyn_pol(indptr, indices, hop, bes, coef_F, coef_omF, psi0, n_timestep,
                     q_points, orb_pos):
    """tbpm_dyn_pol (tbpm.f90:355-493) on caller states psi0 (S, n): per
    q-point, psi1 = (1-F) (density(-q) psi0), psi2 = F psi0, corr(t) =
    Im <psi2 | density(q) psi1> for t = 1 (no step) .. T, both vectors one
    step further each time (density_coef, funcs.f90:125-149).  Returns
    (S, n_q, T)."""
    psi0 = np.atleast_2d(np.asarray(psi0, dtype=np.complex128))
    q = np.atleast_2d(np.asarray(q_points, dtype=np.float64))
    pos = np.asarray(orb_pos, dtype=np.float64)
    out = np.zeros((psi0.shape[0], len(q), n_timestep))
    for iq, qq in enumerate(q):
        pw = qq[0] * pos[:, 0] + qq[1] * pos[:, 1] + qq[2] * pos[:, 2]
        rho = np.cos(pw) + 1j * np.sin(pw)
        for s, w0 in enumerate(psi0):
            psi1 = orc_cheb_apply(indptr, indices, hop, coef_omF, np.conj(rho) * w0)
            psi2 = orc_cheb_apply(indptr, indices, hop, coef_F, w0)
            out[s, iq, 0] = np.vdot(psi2, rho * psi1).imag
            for t in range(1, n_timestep):
                psi1 = orc_step(indptr, indices, hop, bes, psi1, True)
                psi2 = orc_step(indptr, indices, hop, bes, psi2, True)
                out[s, iq, t] = np.vdot(psi2, rho * psi1).imag
    return out


def jackson_kernel(n_kernel):
    """jackson_kernel (kpm.f90:8-29)."""
    g = np.zeros(n_kernel)
    g[0] = 0.5
    q = math.pi / n_kernel
    cotq = math.cos(q) / math.sin(q)
    for k in range(1, n_kernel):
        g[k] = ((n_kernel - k) * math.cos(q * k) + cotq * math.sin(q * k)) / n_kernel
    return g


def orc_tbpm_kbdc(indptr, indices, hop, dx, dy, rescale, psi0, n_kernel, direction=1):
    """tbpm_kbdc (tbpm.f90:836-960) on caller states psi0 (S, n): per state
    g_i g_j <T_i(H) J_x psi | J_b T_j(H) psi> (S, N, N), [s, i, j]; the
    currents J_a = i rescale hop d_a on H's structure (funcs.f90:12-33),
    T_j by the recurrence of tbpm.f90:892-900 with orc_csr_mv."""
    psi0 = np.atleast_2d(np.asarray(psi0, dtype=np.complex128))
    hop = np.asarray(hop, dtype=np.complex128)
    jx = 1j * rescale * hop * np.asarray(dx)
    jb = jx if direction == 1 else 1j * rescale * hop * np.asarray(dy)
    g = jackson_kernel(n_kernel)
    out = np.zeros((psi0.shape[0], n_kernel, n_kernel), dtype=np.complex128)

    def cheb_vectors(v):
        t = [v]
        if n_kernel > 1:
            t.append(orc_csr_mv(indptr, indices, hop, v))
        for j in range(2, n_kernel):
            t.append(2.0 * orc_csr_mv(indptr, indices, hop, t[j - 1]) - t[j - 2])
        return t

    for s, w in enumerate(psi0):
        W = np.array([orc_csr_mv(indptr, indices, jb, t) for t in cheb_vectors(w)])
        P = np.array(cheb_vectors(orc_csr_mv(indptr, indices, jx, w)))
        out[s] = (g[:, None] * g[None, :]) * (P.conj() @ W.T)
    return out


def orc_cond_from_trace(mu_mn, mu, h_rescale, beta, ne_integral):
    """cond_from_trace (analysis.f90:6-69), loops as written (small sizes):
    Gamma_mn of get_gamma_mn (kpm.f90:31-56), Fermi_dist (funcs.f90:36-57)."""
    N = mu_mn.shape[0]
    dE = math.pi / ne_integral
    ne = ne_integral - 1
    sums = np.zeros(ne)
    for k in range(1, ne + 1):
        x = k * dE
        sx, cx = math.sin(x), math.cos(x)
        acc = 0.0
        for j in range(N):
            for i in range(N):
                gam = (complex(cx, -j * sx) * math.cos(i * x) * complex(math.cos(j * x),
                                                                         math.sin(j * x))
                       + complex(cx, i * sx) * math.cos(j * x) * complex(math.cos(i * x),
                                                                         -math.sin(i * x)))
                acc += (gam * mu_mn[i, j]).real
        sums[k - 1] = acc
    out = np.zeros(len(mu))
    for i, m in enumerate(mu):
        dcx = 0.0
        for k in range(1, ne + 1):
            x = k * dE
            en = math.cos(x) * h_rescale
            if en >= m:
                e = math.exp(beta * (m - en))
                fd = e / (1 + e)
            else:
                e = math.exp(beta * (en - m))
                fd = 1 / (1 + e)
            dcx += sums[k - 1] * fd * dE / math.sin(x) ** 3
        out[i] = dcx
    return out


def orc_haydock_coef(indptr, indices, hop, rescale, psi, depth):
    """Haydock_coef (propagation.f90:85-155) with orc_csr_mv: (a, b)."""
    hop = np.asarray(hop, dtype=np.complex128)
    cur = np.asarray(psi, dtype=np.complex128).copy()
    prev = np.zeros_like(cur)
    a = np.zeros(depth, dtype=np.complex128)
    b = np.zeros(depth)
    bprev = 0.0
    for i in range(depth):
        buf = rescale * orc_csr_mv(indptr, indices, hop, cur)
        a[i] = np.vdot(cur, buf)
        buf = a[i] * -1 * cur + (-bprev) * prev + buf
        b[i] = math.sqrt(np.vdot(buf, buf).real)
        buf = buf / b[i]
        prev, cur, bprev = cur, buf, b[i]
    return a, b


def orc_green_function(energy, delta, coefa, coefb):
    """green_function (funcs.f90:151-172) at one energy."""
    z = complex(energy, delta)
    g = 0j
    for i in range(len(coefa) - 1, -1, -1):
        g = 1.0 / (z - coefa[i] - coefb[i] ** 2 * g)
    return g


def orc_csr_mv(indptr, indices, val, x):
    """csr_mv (csr.F90:71-101)."""
    lib = oracle_lib()
    indptr, indices, val = _csr_args(indptr, indices, val)
    x = np.ascontiguousarray(x, dtype=np.complex128)
    out = np.zeros_like(x)
    lib.orc_csr_mv(ctypes.c_int64(len(x)), _ptr(indptr, ctypes.c_int64),
                   _ptr(indices, ctypes.c_int64), _ptr(val), _ptr(x),
                   _ptr(out))
    return out


def orc_fermi_coef(nr_fermi, beta_re, mu_re, one_minus, eps):
    """get_Fermi_cheb_coef (funcs.f90:60-123); returns the kept coefficients."""
    lib = oracle_lib()
    coef = np.zeros(nr_fermi, dtype=np.float64)
    nc = ctypes.c_int64(0)
    rc = lib.orc_fermi_coef(ctypes.c_int64(nr_fermi), ctypes.c_double(beta_re),
                            ctypes.c_double(mu_re),
                            ctypes.c_int(1 if one_minus else 0),
                            ctypes.c_double(eps), _ptr(coef), ctypes.byref(nc))
    if rc:
        raise RuntimeError(f"orc_fermi_coef returned {rc}")
    return coef[:nc.value].copy()


# ---------------------------------------------------------------------------
# The reference Fortran, compiled from tbplas/fortran (oracle/_ref)
# ---------------------------------------------------------------------------
def _fstr(s):
    b = s.encode()
    return ctypes.c_char_p(b), ctypes.c_int64(len(b))


def ref_random_state(n, iseed):
    """random_state (random.f90:36-69) as compiled here (flang intrinsic RNG).

    tbpm_* call it with ``iseed = seed * (i_sample + rank * n_ran_samples)``.
    """
    lib = ref_lib()
    wf = np.zeros(n, dtype=np.complex128)
    lib._QMrandomPrandom_state(_ptr(wf), ctypes.byref(ctypes.c_int(n)),
                               ctypes.byref(ctypes.c_int(iseed)))
    return wf


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def ref_tbpm_dos(bes, indptr, indices, hop, seed, n_timestep, n_samples,
                 chk_step=128, chk_thr=1e-9):
    """The reference tbpm_dos (tbpm.f90:6-93), sample-averaged (T+1,)."""
    lib = ref_lib()
    bes = np.ascontiguousarray(bes, dtype=np.float64)
    ip, ix = _i32(indptr), _i32(indices)
    hop = np.ascontiguousarray(hop, dtype=np.complex128)
    corr = np.zeros(n_timestep + 1, dtype=np.complex128)
    fn, fl = _fstr("unused")
    I = ctypes.c_int
    lib.tbpm_dos_(_ptr(bes), ctypes.byref(I(len(bes))), _ptr(ip, I),
                  ctypes.byref(I(len(ip))), _ptr(ix, I),
                  ctypes.byref(I(len(ix))), _ptr(hop),
                  ctypes.byref(I(len(hop))), ctypes.byref(I(seed)),
                  ctypes.byref(I(n_timestep)), ctypes.byref(I(n_samples)),
                  fn, _ptr(corr), ctypes.byref(I(0)),
                  ctypes.byref(I(chk_step)),
                  ctypes.byref(ctypes.c_double(chk_thr)), fl)
    return corr


def ref_tbpm_accond(bes, beta_re, mu_re, indptr, indices, hop, rescale, dx, dy,
                    seed, n_timestep, n_samples, nr_fermi=2 ** 15,
                    fermi_prec=1e-12, chk_step=128, chk_thr=1e-9):
    """The reference tbpm_accond (tbpm.f90:193-351), averaged (4, T)."""
    lib = ref_lib()
    bes = np.ascontiguousarray(bes, dtype=np.float64)
    ip, ix = _i32(indptr), _i32(indices)
    hop = np.ascontiguousarray(hop, dtype=np.complex128)
    dx = np.ascontiguousarray(dx, dtype=np.float64)
    dy = np.ascontiguousarray(dy, dtype=np.float64)
    corr = np.zeros((4, n_timestep), dtype=np.complex128, order="F")
    fn, fl = _fstr("unused")
    I, D = ctypes.c_int, ctypes.c_double
    r = ctypes.byref
    lib.tbpm_accond_(_ptr(bes), r(I(len(bes))), r(D(beta_re)), r(D(mu_re)),
                     _ptr(ip, I), r(I(len(ip))), _ptr(ix, I), r(I(len(ix))),
                     _ptr(hop), r(I(len(hop))), r(D(rescale)), _ptr(dx),
                     r(I(len(dx))), _ptr(dy), r(I(len(dy))), r(I(seed)),
                     r(I(n_timestep)), r(I(n_samples)), r(I(nr_fermi)),
                     r(D(fermi_prec)), fn, _ptr(corr), r(I(0)), r(I(chk_step)),
                     r(D(chk_thr)), fl)
    return np.ascontiguousarray(corr)


def ref_tbpm_dccond(bes, beta_re, mu_re, indptr, indices, hop, rescale, dx, dy,
                    seed, n_timestep, n_samples, t_step, energies, en_inds,
                    chk_step=128, chk_thr=1e-9):
    """The reference tbpm_dccond (tbpm.f90:497-718): (T,), (2, nE_grid, T)."""
    lib = ref_lib()
    bes = np.ascontiguousarray(bes, dtype=np.float64)
    ip, ix = _i32(indptr), _i32(indices)
    hop = np.ascontiguousarray(hop, dtype=np.complex128)
    dx = np.ascontiguousarray(dx, dtype=np.float64)
    dy = np.ascontiguousarray(dy, dtype=np.float64)
    energies = np.ascontiguousarray(energies, dtype=np.float64)
    en = _i32(en_inds)
    dos = np.zeros(n_timestep, dtype=np.complex128)
    dc = np.zeros((2, len(energies), n_timestep), dtype=np.complex128,
                  order="F")
    f1, l1 = _fstr("unused1")
    f2, l2 = _fstr("unused2")
    I, D = ctypes.c_int, ctypes.c_double
    r = ctypes.byref
    lib.tbpm_dccond_(_ptr(bes), r(I(len(bes))), r(D(beta_re)), r(D(mu_re)),
                     _ptr(ip, I), r(I(len(ip))), _ptr(ix, I), r(I(len(ix))),
                     _ptr(hop), r(I(len(hop))), r(D(rescale)), _ptr(dx),
                     r(I(len(dx))), _ptr(dy), r(I(len(dy))), r(I(seed)),
                     r(I(n_timestep)), r(I(n_samples)), r(D(t_step)),
                     _ptr(energies), r(I(len(energies))), _ptr(en, I),
                     r(I(len(en))), f1, f2, _ptr(dos), _ptr(dc), r(I(0)),
                     r(I(chk_step)), r(D(chk_thr)), l1, l2)
    return dos, np.ascontiguousarray(dc)


def ref_tbpm_ldos(site_indices, bes, indptr, indices, hop, seed, n_timestep, n_samples,
                  chk_step=128, chk_thr=1e-9):
    """The reference tbpm_ldos (tbpm.f90:97-189), sample-averaged (T+1,)."""
    lib = ref_lib()
    si = _i32(site_indices)
    bes = np.ascontiguousarray(bes, dtype=np.float64)
    ip, ix = _i32(indptr), _i32(indices)
    hop = np.ascontiguousarray(hop, dtype=np.complex128)
    corr = np.zeros(n_timestep + 1, dtype=np.complex128)
    fn, fl = _fstr("unused")
    I, D = ctypes.c_int, ctypes.c_double
    r = ctypes.byref
    lib.tbpm_ldos_(_ptr(si, I), r(I(len(si))), _ptr(bes), r(I(len(bes))), _ptr(ip, I),
                   r(I(len(ip))), _ptr(ix, I), r(I(len(ix))), _ptr(hop), r(I(len(hop))),
                   r(I(seed)), r(I(n_timestep)), r(I(n_samples)), fn, _ptr(corr), r(I(0)),
                   r(I(chk_step)), r(D(chk_thr)), fl)
    return corr


def ref_tbpm_psi_t(bes, indptr, indices, hop, psi0, time_log, n_timestep, chk_step=128,
                   chk_thr=1e-9):
    """The reference tbpm_psi_t (tbpm.f90:1029-1093): (n_log, n)."""
    lib = ref_lib()
    bes = np.ascontiguousarray(bes, dtype=np.float64)
    ip, ix = _i32(indptr), _i32(indices)
    hop = np.ascontiguousarray(hop, dtype=np.complex128)
    psi0 = np.ascontiguousarray(psi0, dtype=np.complex128)
    tl = _i32(time_log)
    out = np.zeros((len(tl), len(psi0)), dtype=np.complex128)   # Fortran (n, n_log)
    I, D = ctypes.c_int, ctypes.c_double
    r = ctypes.byref
    lib.tbpm_psi_t_(_ptr(bes), r(I(len(bes))), _ptr(ip, I), r(I(len(ip))), _ptr(ix, I),
                    r(I(len(ix))), _ptr(hop), r(I(len(hop))), _ptr(psi0), r(I(len(psi0))),
                    _ptr(tl, I), r(I(len(tl))), r(I(n_timestep)), r(I(chk_step)),
                    r(D(chk_thr)), _ptr(out))
    return out


def ref_tbpm_eigenstates(bes, indptr, indices, hop, seed, n_timestep, n_samples, t_step,
                         energies, chk_step=128, chk_thr=1e-9):
    """The reference tbpm_eigenstates (tbpm.f90:721-830): (n_E, n)."""
    lib = ref_lib()
    bes = np.ascontiguousarray(bes, dtype=np.float64)
    ip, ix = _i32(indptr), _i32(indices)
    hop = np.ascontiguousarray(hop, dtype=np.complex128)
    en = np.ascontiguousarray(energies, dtype=np.float64)
    n = len(ip) - 1
    out = np.zeros((n, len(en)))                                  # Fortran (n_E, n)
    I, D = ctypes.c_int, ctypes.c_double
    r = ctypes.byref
    lib.tbpm_eigenstates_(_ptr(bes), r(I(len(bes))), _ptr(ip, I), r(I(len(ip))), _ptr(ix, I),
                          r(I(len(ix))), _ptr(hop), r(I(len(hop))), r(I(seed)),
                          r(I(n_timestep)), r(I(n_samples)), r(D(t_step)), _ptr(en),
                          r(I(len(en))), _ptr(out), r(I(0)), r(I(chk_step)), r(D(chk_thr)))
    return np.ascontiguousarray(out.T)


def ref_tbpm_dyn_pol(bes, beta_re, mu_re, indptr, indices, hop, rescale, dx, dy, orb_pos,
                     seed, n_timestep, n_samples, q_points, nr_fermi=2 ** 15, fermi_prec=1e-12,
                     chk_step=128, chk_thr=1e-9):
    """The reference tbpm_dyn_pol (tbpm.f90:355-493): (n_q, T)."""
    lib = ref_lib()
    bes = np.ascontiguousarray(bes, dtype=np.float64)
    ip, ix = _i32(indptr), _i32(indices)
    hop = np.ascontiguousarray(hop, dtype=np.complex128)
    dx = np.ascontiguousarray(dx, dtype=np.float64)
    dy = np.ascontiguousarray(dy, dtype=np.float64)
    pos = np.asarray(orb_pos, dtype=np.float64)
    sx, sy, sz = (np.ascontiguousarray(pos[:, i]) for i in range(3))
    q = np.atleast_2d(np.asarray(q_points, dtype=np.float64))
    qf = np.ascontiguousarray(q.T)                      # Fortran (n_q, 3)
    corr = np.zeros((n_timestep, len(q)))               # Fortran (n_q, T)
    fn, fl = _fstr("unused")
    I, D = ctypes.c_int, ctypes.c_double
    r = ctypes.byref
    lib.tbpm_dyn_pol_(_ptr(bes), r(I(len(bes))), r(D(beta_re)), r(D(mu_re)), _ptr(ip, I),
                      r(I(len(ip))), _ptr(ix, I), r(I(len(ix))), _ptr(hop), r(I(len(hop))),
                      r(D(rescale)), _ptr(dx), r(I(len(dx))), _ptr(dy), r(I(len(dy))),
                      _ptr(sx), r(I(len(sx))), _ptr(sy), r(I(len(sy))), _ptr(sz), r(I(len(sz))),
                      r(I(seed)), r(I(n_timestep)), r(I(n_samples)), r(I(nr_fermi)),
                      r(D(fermi_prec)), _ptr(qf), r(I(len(q))), fn, _ptr(corr), r(I(0)),
                      r(I(chk_step)), r(D(chk_thr)), fl)
    return np.ascontiguousarray(corr.T)


def ref_tbpm_kbdc(seed, indptr, indices, hop, rescale, dx, dy, n_samples, n_kernel,
                  direction=1):
    """The reference tbpm_kbdc (tbpm.f90:836-960): (N, N) [i, j]."""
    lib = ref_lib()
    ip, ix = _i32(indptr), _i32(indices)
    hop = np.ascontiguousarray(hop, dtype=np.complex128)
    dx = np.ascontiguousarray(dx, dtype=np.float64)
    dy = np.ascontiguousarray(dy, dtype=np.float64)
    out = np.zeros((n_kernel, n_kernel), dtype=np.complex128)   # Fortran (i, j) -> [j, i]
    I, D = ctypes.c_int, ctypes.c_double
    r = ctypes.byref
    lib.tbpm_kbdc_(r(I(seed)), _ptr(ip, I), r(I(len(ip))), _ptr(ix, I), r(I(len(ix))),
                   _ptr(hop), r(I(len(hop))), r(D(rescale)), _ptr(dx), r(I(len(dx))),
                   _ptr(dy), r(I(len(dy))), r(I(n_samples)), r(I(n_kernel)),
                   r(I(direction)), _ptr(out), r(I(0)))
    return np.ascontiguousarray(out.T)


def ref_cond_from_trace(mu_mn, mu, h_rescale, beta, ne_integral, fermi_prec=1e-12):
    """The reference cond_from_trace (analysis.f90:6-69)."""
    lib = ref_lib()
    m = np.asfortranarray(mu_mn, dtype=np.complex128)
    mu = np.ascontiguousarray(mu, dtype=np.float64)
    cond = np.zeros(len(mu))
    I, D = ctypes.c_int, ctypes.c_double
    r = ctypes.byref
    lib.cond_from_trace_(m.ctypes.data_as(ctypes.c_void_p), r(I(m.shape[0])), _ptr(mu),
                         r(I(len(mu))), r(D(h_rescale)), r(D(beta)), r(I(ne_integral)),
                         r(D(fermi_prec)), r(I(0)), _ptr(cond))
    return cond


def ref_ldos_haydock(site_indices, delta, e_range, indptr, indices, hop, rescale, seed, depth,
                     n_timestep, n_samples):
    """The reference ldos_haydock (tbpm.f90:964-1024): (energy, ldos), each
    (2T+1,)."""
    lib = ref_lib()
    si = _i32(site_indices)
    ip, ix = _i32(indptr), _i32(indices)
    hop = np.ascontiguousarray(hop, dtype=np.complex128)
    energy = np.zeros(2 * n_timestep + 1)
    ldos = np.zeros(2 * n_timestep + 1)
    fn, fl = _fstr("unused")
    I, D = ctypes.c_int, ctypes.c_double
    r = ctypes.byref
    lib.ldos_haydock_(_ptr(si, I), r(I(len(si))), r(D(delta)), r(D(e_range)), _ptr(ip, I),
                      r(I(len(ip))), _ptr(ix, I), r(I(len(ix))), _ptr(hop), r(I(len(hop))),
                      r(D(rescale)), r(I(seed)), r(I(depth)), r(I(n_timestep)),
                      r(I(n_samples)), fn, _ptr(energy), _ptr(ldos), fl)
    return energy, ldos


def ref_fermi_coef(nr_fermi, beta_re, mu_re, one_minus, eps):
    """The reference get_Fermi_cheb_coef (funcs.f90:60-123)."""
    lib = ref_lib()
    coef = np.zeros(nr_fermi + 1, dtype=np.float64)
    nc = ctypes.c_int(0)
    fn = getattr(lib, "_QMfuncsPget_fermi_cheb_coef")
    fn(_ptr(coef), ctypes.byref(nc), ctypes.byref(ctypes.c_int(nr_fermi)),
       ctypes.byref(ctypes.c_double(beta_re)),
       ctypes.byref(ctypes.c_double(mu_re)),
       ctypes.byref(ctypes.c_int(1 if one_minus else 0)),
       ctypes.byref(ctypes.c_double(eps)))
    return coef[:nc.value].copy()


# ---------------------------------------------------------------------------
# Host-side restatements
# ---------------------------------------------------------------------------
def bessel_series(time_scaled, precision=1e-14, bessel_max=250):
    """Solver._get_bessel_series (solver.py:194-224) via scipy.special.jv.

    ``time_scaled = time_step * rescale``.  Keeps J_i while |J_i| > precision
    or |J_{i+1}| > precision; raises if Bessel_max is reached first.
    """
    import scipy.special as spec
    out = []
    for i in range(bessel_max):
        v = spec.jv(i, time_scaled)
        if abs(v) > precision:
            out.append(float(v))
        else:
            if abs(spec.jv(i + 1, time_scaled)) > precision:
                out.append(float(v))
            else:
                return out
    raise ValueError("Bessel_max too low")


def build_ham_csr(orb_eng, hop_i, hop_j, hop_v):
    """Sample.build_ham_csr (sample.py:375-391): scipy dia + csr + csr^H.

    scipy is the reference's own third-party dependency here; its canonical
    form (sorted columns, duplicates summed, explicit zeros dropped) is the
    contract.  Returns (indptr, indices, data) as int64/int64/complex128.
    """
    from scipy.sparse import dia_matrix, csr_matrix
    n = len(orb_eng)
    shape = (n, n)
    ham_dia = dia_matrix((orb_eng, 0), shape=shape)
    ham_half = csr_matrix((hop_v, (hop_i, hop_j)), shape=shape)
    ham = ham_dia + ham_half + ham_half.getH()
    return (ham.indptr.astype(np.int64), ham.indices.astype(np.int64),
            ham.data.astype(np.complex128))


def build_ham_dxy_loop(orb_eng, hop_i, hop_j, hop_v, dr, cutoff=1e-5):
    """build_ham_dxy_fast (sample.pyx:148-257), loop for loop (small inputs)."""
    n, m = len(orb_eng), len(hop_i)
    indptr = np.zeros(n + 1, dtype=np.int64)
    for io in range(n):
        if abs(orb_eng[io]) >= cutoff:
            indptr[io + 1] += 1
    for ih in range(m):
        indptr[hop_i[ih] + 1] += 1
        indptr[hop_j[ih] + 1] += 1
    for io in range(1, n + 1):
        indptr[io] = indptr[io - 1] + indptr[io]
    nnz = indptr[n]
    indices = np.zeros(nnz, dtype=np.int64)
    hop = np.zeros(nnz, dtype=np.complex128)
    dx = np.zeros(nnz, dtype=np.float64)
    dy = np.zeros(nnz, dtype=np.float64)
    cnt = np.zeros(n, dtype=np.int64)

    def free_ptr(i):  # _get_free_ptr (sample.pyx:117-145)
        p = indptr[i] + cnt[i]
        cnt[i] += 1
        return p

    for io in range(n):
        if abs(orb_eng[io]) >= cutoff:
            p = free_ptr(io)
            indices[p] = io
            hop[p] = orb_eng[io]
    for ih in range(m):
        ii, jj = hop_i[ih], hop_j[ih]
        p = free_ptr(ii)
        indices[p] = jj
        hop[p] = hop_v[ih]
        dx[p] = dr[ih, 0]
        dy[p] = dr[ih, 1]
        p = free_ptr(jj)
        indices[p] = ii
        hop[p] = np.conj(hop_v[ih])
        dx[p] = -dr[ih, 0]
        dy[p] = -dr[ih, 1]
    return indptr, indices, hop, dx, dy


def build_ham_dxy_stable(orb_eng, hop_i, hop_j, hop_v, dr, cutoff=1e-5):
    """Same order as build_ham_dxy_loop, vectorised with a stable sort.

    The reference fills each row in arrival order: every kept diagonal first
    (row io), then for each hop ih the entry (ii, jj) followed by (jj, ii).
    A stable sort by row of that emission stream reproduces it exactly.
    Pinned against build_ham_dxy_loop on small inputs.
    """
    n = len(orb_eng)
    hop_i = np.asarray(hop_i, dtype=np.int64)
    hop_j = np.asarray(hop_j, dtype=np.int64)
    keep = np.nonzero(np.abs(orb_eng) >= cutoff)[0].astype(np.int64)
    m = len(hop_i)
    rows = np.empty(len(keep) + 2 * m, dtype=np.int64)
    cols = np.empty_like(rows)
    vals = np.empty(len(rows), dtype=np.complex128)
    dxs = np.zeros(len(rows))
    dys = np.zeros(len(rows))
    k = len(keep)
    rows[:k] = keep
    cols[:k] = keep
    vals[:k] = np.asarray(orb_eng)[keep]
    rows[k::2], rows[k + 1::2] = hop_i, hop_j
    cols[k::2], cols[k + 1::2] = hop_j, hop_i
    vals[k::2], vals[k + 1::2] = hop_v, np.conj(hop_v)
    dxs[k::2], dxs[k + 1::2] = dr[:, 0], -dr[:, 0]
    dys[k::2], dys[k + 1::2] = dr[:, 1], -dr[:, 1]
    order = np.argsort(rows, kind="stable")
    indptr = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(np.bincount(rows, minlength=n), out=indptr[1:])
    return indptr, cols[order], vals[order], dxs[order], dys[order]


def graphene_hops_loop(L1, L2, t=-2.7, pbc=(True, True), a=0.246):
    """Graphene supercell hop list, loop for loop.

    Primitive cell: make_graphene_diamond (materials/graphene.py:16-31),
    lattice a=0.246 nm at gamma=60 deg, orbitals at (0,0) and (1/3,1/3).
    Reduced hop order from check_conj (builder/base.py:76-92) on the
    insertion order of add_hopping: [(0,0,0,0,1), (1,0,0,1,0), (0,1,0,1,0)].
    Supercell order: build_hop_pbc (super.pyx:860-986) loops ia, ib, ic, ih;
    orbital ids from _id_pc2sc (super.pyx:19-43); positions build_orb_pos
    (super.pyx:569-612); dr includes the lattice shift of wrapped cells
    (_fast_div, super.pyx:496-539).  Open directions drop hops leaving the
    cell (split_pc_hop / build_hop_gen, super.pyx:643-857), which for this
    cell keeps the same relative order.
    Returns hop_i, hop_j, hop_v, dr, orb_pos.
    """
    a1 = np.array([a, 0.0, 0.0])
    gam = 60.0 / 180 * math.pi     # gen_lattice_vectors (base/lattice.py:36-41)
    a2 = np.array([a * math.cos(gam), a * math.sin(gam), 0.0])
    a3 = np.array([0.0, 0.0, 1.0])
    lat = np.array([a1, a2, a3])
    frac = np.array([[0.0, 0.0, 0.0], [1 / 3., 1 / 3., 0.0]])
    pc_hops = [(0, 0, 0, 0, 1), (1, 0, 0, 1, 0), (0, 1, 0, 1, 0)]
    dim = (L1, L2, 1)
    n = 2 * L1 * L2
    orb_pos = np.zeros((n, 3))
    for ia in range(L1):
        for ib in range(L2):
            for io in range(2):
                idx = (ia * L2 + ib) * 2 + io
                ra, rb, rc = frac[io, 0] + ia, frac[io, 1] + ib, frac[io, 2]
                orb_pos[idx] = ra * lat[0] + rb * lat[1] + rc * lat[2]
    sc_lat = np.array([lat[0] * L1, lat[1] * L2, lat[2]])
    hi, hj, dr = [], [], []
    per = (bool(pbc[0]), bool(pbc[1]))
    # split_pc_hop (super.pyx:786-857): a hop is "free" if it crosses an open
    # direction; periodic hops are emitted first by build_hop_pbc, free ones
    # afterwards by build_hop_gen (super.py:703-741).
    hp = [h for h in pc_hops if all(per[d] or h[d] == 0 for d in range(2))]
    hf = [h for h in pc_hops if h not in hp]

    def emit(ia, ib, h):
        ra, rb, rc, oi, oj = h
        ja0, jb0 = ia + ra, ib + rb
        if (not per[0] and not 0 <= ja0 < L1) or \
                (not per[1] and not 0 <= jb0 < L2):
            return
        ja, jb = ja0 % L1, jb0 % L2
        i = (ia * L2 + ib) * 2 + oi
        j = (ja * L2 + jb) * 2 + oj
        na, nb = ja0 // L1, jb0 // L2
        hi.append(i)
        hj.append(j)
        dr.append(orb_pos[j] - orb_pos[i] + na * sc_lat[0] + nb * sc_lat[1])

    for ia in range(L1):                 # build_hop_pbc: ia, ib, (ic), ih
        for ib in range(L2):
            for h in hp:
                emit(ia, ib, h)
    for io in range(n):                  # build_hop_gen: io, ih (same orb_i)
        ia, ib, oo = io // (2 * L2), (io // 2) % L2, io % 2
        for h in hf:
            if h[3] == oo:
                emit(ia, ib, h)
    m = len(hi)
    return (np.array(hi, dtype=np.int64), np.array(hj, dtype=np.int64),
            np.full(m, t, dtype=np.complex128), np.array(dr).reshape(m, 3),
            orb_pos)


def orc_supercell_hops(pc_lattice, pc_orb_frac, pc_hop_ind, pc_hop_eng, dim, pbc,
                       vacancies=()):
    """SuperCell.get_hop without a hop modifier (builder/super.py:703-775),
    loop for loop: orbitals (ia, ib, ic, orb) minus vacancies
    (super.pyx:383-443), positions build_orb_pos (super.pyx:569-612);
    without vacancies split_pc_hop (786-857) then build_hop_pbc (860-986)
    and build_hop_gen (643-783), with vacancies build_hop_gen on all terms;
    _wrap_pbc, _fast_div and _id_pc2sc(_vac2) (19-150, 446-539).  Returns
    orb_pos, hop_i, hop_j, hop_v, dr."""
    lat = np.asarray(pc_lattice, dtype=np.float64)
    frac = np.asarray(pc_orb_frac, dtype=np.float64)
    d0, d1, d2 = (int(x) for x in dim)
    npc = len(frac)
    per = [bool(x) for x in pbc]
    vac = sorted({((a * d1 + b) * d2 + c) * npc + o for a, b, c, o in vacancies})
    vset = set(vac)
    orb_id, pos = [], []
    for ia in range(d0):
        for ib in range(d1):
            for ic in range(d2):
                for o in range(npc):
                    if (((ia * d1 + ib) * d2 + ic) * npc + o) in vset:
                        continue
                    orb_id.append((ia, ib, ic, o))
                    ra, rb, rc = frac[o, 0] + ia, frac[o, 1] + ib, frac[o, 2] + ic
                    pos.append([ra * lat[0, x] + rb * lat[1, x] + rc * lat[2, x]
                                for x in range(3)])
    pos = np.array(pos).reshape(-1, 3)
    sc = lat * np.array([d0, d1, d2], dtype=np.float64)[:, None]
    dims = (d0, d1, d2)

    def wrap(j, n, p):
        if not p:
            return j if 0 <= j < n else -1
        return j % n

    def pc2sc(a, b, c, o):
        return ((a * d1 + b) * d2 + c) * npc + o

    def sc_vac(idx):
        if idx in vset:
            return -1
        return idx - sum(1 for v in vac if v < idx)

    def dr_of(i, j, n):
        return [pos[j, x] - pos[i, x] + n[0] * sc[0, x] + n[1] * sc[1, x] + n[2] * sc[2, x]
                for x in range(3)]

    hops = [tuple(int(v) for v in h) for h in np.asarray(pc_hop_ind).reshape(-1, 5)]
    eng = list(np.asarray(pc_hop_eng, dtype=np.complex128).reshape(-1))
    hi, hj, hv, dr = [], [], [], []
    if not vac:
        hp = [k for k, h in enumerate(hops) if all(per[x] or h[x] == 0 for x in range(3))]
    else:
        hp = []
    hf = [k for k in range(len(hops)) if k not in hp]
    for ia in range(d0):                                  # build_hop_pbc
        for ib in range(d1):
            for ic in range(d2):
                oi = pc2sc(ia, ib, ic, 0)
                for k in hp:
                    h = hops[k]
                    if h[0] == 0 and h[1] == 0 and h[2] == 0:
                        oj, n = oi, (0, 0, 0)
                    else:
                        j0 = (ia + h[0], ib + h[1], ic + h[2])
                        oj = pc2sc(*(wrap(j0[x], dims[x], per[x]) for x in range(3)), 0)
                        n = tuple(j0[x] // dims[x] for x in range(3))
                    i, j = oi + h[3], oj + h[4]
                    hi.append(i)
                    hj.append(j)
                    hv.append(eng[k])
                    dr.append(dr_of(i, j, n))
    for io, (ia, ib, ic, o) in enumerate(orb_id):         # build_hop_gen
        for k in hf:
            h = hops[k]
            if h[3] != o:
                continue
            j0 = (ia + h[0], ib + h[1], ic + h[2])
            jw = [wrap(j0[x], dims[x], per[x]) for x in range(3)]
            if -1 in jw:
                continue
            j = pc2sc(jw[0], jw[1], jw[2], h[4])
            if vac:
                j = sc_vac(j)
                if j == -1:
                    continue
            hi.append(io)
            hj.append(j)
            hv.append(eng[k])
            dr.append(dr_of(io, j, tuple(j0[x] // dims[x] for x in range(3))))
    m = len(hi)
    return (pos, np.array(hi, dtype=np.int64), np.array(hj, dtype=np.int64),
            np.array(hv, dtype=np.complex128), np.array(dr).reshape(m, 3))


def graphene_hops_vec(L1, L2, t=-2.7, pbc=(True, True), a=0.246):
    """graphene_hops_loop vectorised with numpy (same order and bits, pinned
    against the loop), for sizes the loop cannot reach (the CPU baseline's
    8.4M-orbital sample).  Returns hop_i, hop_j, hop_v, dr, orb_pos."""
    L1, L2 = int(L1), int(L2)
    a1 = np.array([a, 0.0, 0.0])
    gam = 60.0 / 180 * math.pi
    a2 = np.array([a * math.cos(gam), a * math.sin(gam), 0.0])
    lat = np.array([a1, a2, [0.0, 0.0, 1.0]])
    n = 2 * L1 * L2
    ia = np.repeat(np.arange(L1), L2)
    ib = np.tile(np.arange(L2), L1)
    pos = np.empty((n, 3))
    for io, (fa, fb) in enumerate(((0.0, 0.0), (1 / 3., 1 / 3.))):
        pos[io::2] = ((fa + ia)[:, None] * lat[0] + (fb + ib)[:, None] * lat[1]
                      + 0.0 * lat[2])
    sc1, sc2 = lat[0] * L1, lat[1] * L2
    cell = ia * L2 + ib
    pc = [(0, 0, 0, 1), (1, 0, 1, 0), (0, 1, 1, 0)]
    per = (bool(pbc[0]), bool(pbc[1]))
    periodic = [h for h in pc if (per[0] or h[0] == 0) and (per[1] or h[1] == 0)]
    free = [h for h in pc if h not in periodic]

    def emit(hops):
        ii, jj, dd, ok = [], [], [], []
        for (ra, rb, oi, oj) in hops:
            ja0, jb0 = ia + ra, ib + rb
            valid = np.ones(L1 * L2, dtype=bool)
            if not per[0]:
                valid &= ja0 < L1
            if not per[1]:
                valid &= jb0 < L2
            na, nb = ja0 // L1, jb0 // L2
            ja, jb = ja0 - na * L1, jb0 - nb * L2
            i = cell * 2 + oi
            j = (ja * L2 + jb) * 2 + oj
            ii.append(i)
            jj.append(j)
            dd.append(pos[j] - pos[i] + na[:, None] * sc1 + nb[:, None] * sc2)
            ok.append(valid)
        if not hops:
            z = np.zeros(0, dtype=np.int64)
            return z, z, np.zeros((0, 3)), np.zeros(0, dtype=bool)
        return (np.stack(ii, 1).reshape(-1), np.stack(jj, 1).reshape(-1),
                np.stack(dd, 1).reshape(-1, 3), np.stack(ok, 1).reshape(-1))

    parts = [emit(periodic), emit(free)]
    I = np.concatenate([p[0] for p in parts])
    J = np.concatenate([p[1] for p in parts])
    D = np.concatenate([p[2] for p in parts])
    V = np.concatenate([p[3] for p in parts])
    I, J, D = I[V], J[V], D[V]
    return (I.astype(np.int64), J.astype(np.int64), np.full(len(I), t, dtype=np.complex128),
            np.ascontiguousarray(D), pos)


def get_rescale_loop(orb_eng, hop_i, hop_j, hop_v):
    """get_rescale (sample.pyx:75-114): max over rows of the signed on-site
    energy plus the moduli of the row's hopping terms, each modulus formed
    as sqrt(re**2 + im**2) (sample.pyx:110), not hypot.

    Pinned by the closed form max_i(orb_eng_i + sum_{j != i} |H_ij|) on a
    dense H with signed on-site energies and complex hoppings, and by a row
    whose negative on-site energy must not count as |orb_eng|."""
    s = np.array(orb_eng, dtype=np.float64).copy()
    for ih in range(len(hop_i)):
        z = complex(hop_v[ih])
        v = math.sqrt(z.real ** 2 + z.imag ** 2)
        s[hop_i[ih]] += v
        s[hop_j[ih]] += v
    return float(np.max(s))


def set_mag_field_loop(hop_i, hop_j, hop_v, dr, orb_pos, intensity, gauge=0):
    """set_mag_field (sample.pyx:18-72), in place on a copy of hop_v.

    Pinned by the flux through every closed loop: around each hexagon of an
    open graphene flake, traversed counter-clockwise, the product of the
    amplitudes H[i_{k+1}, i_k] carries the phase 2 pi B A_hex / Phi_0
    (Phi_0 = h/e = 4135.666734 T nm^2) in each of the three gauges, and the
    spectrum of the flake is the same in all three (a gauge change is a
    diagonal unitary)."""
    factor = math.pi / 4135.666734
    out = np.array(hop_v, dtype=np.complex128).copy()
    for ih in range(len(hop_i)):
        ii, jj = hop_i[ih], hop_j[ih]
        if gauge == 0:
            phase = factor * intensity * dr[ih, 0] * (orb_pos[jj, 1] + orb_pos[ii, 1])
        elif gauge == 1:
            phase = -factor * intensity * dr[ih, 1] * (orb_pos[jj, 0] + orb_pos[ii, 0])
        else:
            sx = orb_pos[jj, 0] + orb_pos[ii, 0]
            sy = orb_pos[jj, 1] + orb_pos[ii, 1]
            phase = factor * 0.5 * intensity * (dr[ih, 0] * sx - dr[ih, 1] * sy)
        out[ih] = out[ih] * (math.cos(phase) + 1j * math.sin(phase))
    return out


# ---------------------------------------------------------------------------
# Exact oracles (independent of the Chebyshev method)
# ---------------------------------------------------------------------------
def kspace_graphene_corr(L1, L2, hop, psi0, tau, steps, fwd=True):
    """Exact <psi0| exp(-i Ht tau k) |psi0> for pristine periodic graphene.

    ``hop`` is the (rescaled) nearest-neighbour hopping, ``tau`` the scaled
    time step (time_step * rescale), orbital order (ia*L2+ib)*2+io as in
    graphene_hops_loop.  H is block diagonal in k with the 2x2 block
    [[0, f], [f*, 0]], f = hop + hop* (e^{-i k1} + e^{-i k2}); the weights of
    the two bands come from one 2-D FFT of each sublattice of psi0.
    ``steps`` is an iterable of integer time indices.
    """
    psi0 = np.asarray(psi0).reshape(L1, L2, 2)
    A = np.fft.fft2(psi0[:, :, 0])
    B = np.fft.fft2(psi0[:, :, 1])
    k1 = 2 * np.pi * np.fft.fftfreq(L1)[:, None]
    k2 = 2 * np.pi * np.fft.fftfreq(L2)[None, :]
    f = hop + np.conj(hop) * (np.exp(-1j * k1) + np.exp(-1j * k2))
    e = np.abs(f)
    ph = np.where(e > 0, f / np.where(e > 0, e, 1), 1.0)
    # eigenvectors (1, +-conj(ph))/sqrt2 for eigenvalues +-e
    up = (A + ph * B) / np.sqrt(2)        # <v+|.> with v+ = (1, conj(ph))/sqrt2
    dn = (A - ph * B) / np.sqrt(2)
    N = L1 * L2
    wu = (np.abs(up) ** 2).ravel() / N
    wd = (np.abs(dn) ** 2).ravel() / N
    e = e.ravel()
    sgn = -1.0 if fwd else 1.0
    out = []
    for k in steps:
        th = sgn * tau * k * e
        out.append(np.sum(wu * np.exp(1j * th)) + np.sum(wd * np.exp(-1j * th)))
    return np.array(out, dtype=np.complex128)


def kspace_graphene_corr_multi(L1, L2, hop, psi0s, tau, steps, fwd=True):
    """kspace_graphene_corr for several states (rows of ``psi0s``) at once:
    the band energies and the phase factors exp(-+i E tau k) are shared, so
    each sampled step costs one complex exponential over the k grid and one
    (states x k) product.  Returns (states, len(steps)) complex128."""
    psi0s = np.asarray(psi0s).reshape(-1, L1, L2, 2)
    S = psi0s.shape[0]
    k1 = 2 * np.pi * np.fft.fftfreq(L1)[:, None]
    k2 = 2 * np.pi * np.fft.fftfreq(L2)[None, :]
    f = hop + np.conj(hop) * (np.exp(-1j * k1) + np.exp(-1j * k2))
    e = np.abs(f)
    ph = np.where(e > 0, f / np.where(e > 0, e, 1), 1.0)
    N = L1 * L2
    wu = np.empty((S, N))
    wd = np.empty((S, N))
    for s in range(S):
        A = np.fft.fft2(psi0s[s, :, :, 0])
        B = np.fft.fft2(psi0s[s, :, :, 1])
        wu[s] = (np.abs(A + ph * B) ** 2).ravel() / (2 * N)
        wd[s] = (np.abs(A - ph * B) ** 2).ravel() / (2 * N)
    e = e.ravel()
    sgn = -1.0 if fwd else 1.0
    out = np.empty((S, len(list(steps))), dtype=np.complex128)
    for c, k in enumerate(steps):
        z = np.exp(1j * (sgn * tau * k) * e)
        out[:, c] = wu @ z + wd @ np.conj(z)
    return out


def cheb_series_values(coef, x):
    """sum_n coef[n] T_n(x) for |x| <= 1 by Clenshaw's recurrence (the
    polynomial that Fermi, propagation.f90:52-82, applies to H~)."""
    x = np.asarray(x, dtype=np.float64)
    b1 = np.zeros_like(x)
    b2 = np.zeros_like(x)
    for c in coef[:0:-1]:
        b1, b2 = c + 2.0 * x * b1 - b2, b1
    return coef[0] + x * b1 - b2


def kspace_graphene_ac(L1, L2, hops, drs, rescale, coef_F, coef_omF, psi0s, tau, steps):
    """Exact tbpm_accond correlations (tbpm.f90:193-351) of pristine periodic
    graphene, per state: (S, 4, len(steps)) for the entries
    <psi2|J_x psi1_x>, <psi2|J_y psi1_x>, <psi2|J_x psi1_y>, <psi2|J_y psi1_y>
    at the listed time indices t >= 2, with psi1_b = (1 - F)(H~) J_b psi0,
    psi2 = F(H~) psi0, both propagated t - 1 steps of exp(-i H~ tau), F the
    Chebyshev series of the Fermi coefficients (evaluated exactly on the
    band energies) and J_a = i rescale H~ d_a (funcs.f90:12-33).

    ``hops`` / ``drs``: the rescaled values and bond vectors of the three
    hops of cell (0, 0) in graphene_hops_loop order: orbital 0 -> 1 in the
    cell, 1 -> 0 of cell +a1, 1 -> 0 of cell +a2.  Orbital order
    (ia*L2+ib)*2+io.  In k the blocks are H~ = [[0, f], [f*, 0]] and
    J_a = [[0, g_a], [g_a*, 0]]; in the band basis v_s = (1, s conj(f/|f|))
    / sqrt2 the correlation is a sum over s, s' of F(s|f|) (1 - F)(s'|f|)
    <v_s|psi0>* <v_s|J_a|v_s'> <v_s'|J_b psi0> e^{i (s - s')|f| tau (t - 1)},
    with 1/N from Parseval (numpy's unnormalised FFT)."""
    h0, h1, h2 = [complex(h) for h in hops]
    d0, d1, d2 = [np.asarray(d, dtype=np.float64) for d in drs]
    psi0s = np.asarray(psi0s).reshape(-1, L1, L2, 2)
    S, N = psi0s.shape[0], L1 * L2
    k1 = 2 * np.pi * np.fft.fftfreq(L1)[:, None]
    k2 = 2 * np.pi * np.fft.fftfreq(L2)[None, :]
    z1, z2 = np.exp(-1j * k1), np.exp(-1j * k2)
    f = (h0 + np.conj(h1) * z1 + np.conj(h2) * z2).ravel()
    g = [(1j * rescale * (h0 * d0[a] - np.conj(h1) * d1[a] * z1
                          - np.conj(h2) * d2[a] * z2)).ravel() for a in (0, 1)]
    e = np.abs(f)
    ph = np.where(e > 0, f / np.where(e > 0, e, 1), 1.0)
    # F and 1 - F on the band energies +-e: Clenshaw once per distinct value
    ev, inv = np.unique(np.round(e, 14), return_inverse=True)
    Fp = cheb_series_values(coef_F, ev)[inv]
    Fm = cheb_series_values(coef_F, -ev)[inv]
    Op = cheb_series_values(coef_omF, ev)[inv]
    Om = cheb_series_values(coef_omF, -ev)[inv]
    Fs = {1: Fp, -1: Fm}
    Os = {1: Op, -1: Om}
    # <v_s|J_a|v_s'> = (s' g_a conj(ph) + s ph conj(g_a)) / 2
    M = {(a, s, sp): 0.5 * (sp * g[a] * np.conj(ph) + s * ph * np.conj(g[a]))
         for a in (0, 1) for s in (1, -1) for sp in (1, -1)}
    steps = list(steps)
    pairs = [(0, 0), (1, 0), (0, 1), (1, 1)]   # (a, b): <psi2|J_a psi1_b>
    # per (state, pair): the time-independent sum and the weights of
    # e^{+-2i|f| tau (t-1)}, stacked so each step is one matrix product
    w0 = np.zeros(S * 4, dtype=np.complex128)
    wp = np.empty((S * 4, N), dtype=np.complex128)
    wm = np.empty((S * 4, N), dtype=np.complex128)
    for i in range(S):
        A = np.fft.fft2(psi0s[i, :, :, 0]).ravel()
        B = np.fft.fft2(psi0s[i, :, :, 1]).ravel()
        al = {s: (A + s * ph * B) / np.sqrt(2) for s in (1, -1)}
        JA = [g[b] * B for b in (0, 1)]           # components of J_b psi0
        JB = [np.conj(g[b]) * A for b in (0, 1)]
        be = {(b, s): (JA[b] + s * ph * JB[b]) / np.sqrt(2) for b in (0, 1) for s in (1, -1)}
        for q, (a, b) in enumerate(pairs):
            w = {(s, sp): Fs[s] * Os[sp] * np.conj(al[s]) * M[(a, s, sp)] * be[(b, sp)]
                 for s in (1, -1) for sp in (1, -1)}
            w0[i * 4 + q] = np.sum(w[(1, 1)] + w[(-1, -1)])
            wp[i * 4 + q] = w[(1, -1)]
            wm[i * 4 + q] = w[(-1, 1)]
    out = np.zeros((S * 4, len(steps)), dtype=np.complex128)
    for c, t in enumerate(steps):
        zt = np.exp(2j * e * tau * (t - 1))
        out[:, c] = (w0 + wp @ zt + wm @ np.conj(zt)) / N
    return out.reshape(S, 4, len(steps))


def dense_corr(indptr, indices, data, psi0, tau, steps):
    """Exact <psi0|exp(-i H tau k)|psi0> by dense eigendecomposition (n<=4000)."""
    n = len(indptr) - 1
    Hd = np.zeros((n, n), dtype=np.complex128)
    for r in range(n):
        for p in range(indptr[r], indptr[r + 1]):
            Hd[r, indices[p]] += data[p]
    w, v = np.linalg.eigh(Hd)
    c = v.conj().T @ np.asarray(psi0)
    wt = np.abs(c) ** 2
    return np.array([np.sum(wt * np.exp(-1j * w * tau * k)) for k in steps])


# ---------------------------------------------------------------------------
# Analyzer restatements (analysis.py), loop for loop
# ---------------------------------------------------------------------------
def window_hanning(i, tnr):
    """analysis.py:20-28"""
    return 0.5 * (1 + math.cos(math.pi * i / tnr))


def window_exp(i, tnr):
    """analysis.py:31-39"""
    return math.exp(-2. * (i / tnr) ** 2)


def analyzer_calc_dos(corr_dos, tnr, en_range, correct_spin=False,
                      window=window_hanning):
    """Analyzer.calc_dos (analysis.py:89-143)."""
    en_step = 0.5 * en_range / tnr
    energies = np.array([0.5 * i * en_range / tnr - en_range / 2.
                         for i in range(tnr * 2)], dtype=float)
    dos = np.zeros(tnr * 2, dtype=float)
    cn = np.empty(tnr * 2, dtype=complex)
    cn[tnr - 1] = corr_dos.item(0)
    cn[2 * tnr - 1] = window(tnr - 1, tnr) * corr_dos.item(tnr)
    for i in range(tnr - 1):
        cn[tnr + i] = window(i, tnr) * corr_dos.item(i + 1)
        cn[tnr - i - 2] = window(i, tnr) * corr_dos.item(i + 1).conjugate()
    cf = np.fft.ifft(cn)
    for i in range(tnr):
        dos[i + tnr] = abs(cf.item(i))
    for i in range(tnr, 2 * tnr):
        dos[i - tnr] = abs(cf.item(i))
    dos = dos / (np.sum(dos) * en_step)
    if correct_spin:
        dos = 2. * dos
    return energies, dos


def analyzer_calc_ac_cond(corr_ac, tnr, en_range, beta, ac_prefactor,
                          correct_spin=False, window=window_exp):
    """Analyzer.calc_ac_cond (analysis.py:161-243)."""
    from scipy.signal import hilbert
    t_step = math.pi / en_range
    omegas = np.array([i * en_range / tnr for i in range(tnr)], dtype=float)
    ac_real = np.zeros((4, tnr), dtype=float)
    for j in range(4):
        for i in range(tnr):
            omega = omegas.item(i)
            acv = 0.
            for k in range(tnr):
                acv += 2. * window(k + 1, tnr) * math.sin(omega * k * t_step) \
                    * corr_ac.item(j, k).imag
            if omega == 0.:
                acv = 0.
            else:
                acv = ac_prefactor * t_step * acv \
                    * (math.exp(-beta * omega) - 1) / omega
            ac_real[j, i] = acv
    ac_imag = np.zeros((4, tnr), dtype=float)
    for j in range(4):
        sigma = np.zeros(2 * tnr, dtype=float)
        for i in range(tnr):
            sigma[tnr + i] = ac_real.item(j, i)
            sigma[tnr - i] = ac_real.item(j, i)
        ac_imag[j, :] = np.imag(hilbert(sigma))[tnr:2 * tnr]
    ac = ac_real + 1j * ac_imag
    if correct_spin:
        ac = 2. * ac
    return omegas, ac


def analyzer_calc_dc_cond(corr_dos, corr_dc, tnr, en_range, energy_limits,
                          dc_prefactor, correct_spin=False,
                          window_dos=window_hanning, window_dc=window_exp):
    """Analyzer.calc_dc_cond (analysis.py:375-458)."""
    corr_dos = np.insert(corr_dos, 0, 1.0)
    energies_dos, dos = analyzer_calc_dos(corr_dos, tnr, en_range, False,
                                          window_dos)
    t_step = 2 * math.pi / en_range
    qe = np.where((energies_dos >= energy_limits[0]) &
                  (energies_dos <= energy_limits[1]))[0]
    energies = energies_dos[qe]
    dc = np.zeros((2, len(qe)))
    for i in range(2):
        for j in range(len(qe)):
            en = energies.item(j)
            dos_val = dos.item(qe.item(j))
            dc_val = 0.
            for k in range(tnr):
                w = window_dc(k + 1, tnr)
                phi = k * t_step * en
                c_exp = math.cos(phi) - 1j * math.sin(phi)
                dc_val += w * (c_exp * corr_dc.item(i, j, k)).real
            dc[i, j] = dc_prefactor * t_step * dos_val * dc_val
    if correct_spin:
        dc = 2. * dc
    return energies, dc


def analyzer_calc_diff_coeff(corr_dc, tnr, en_range, energy_limits, window_dc=window_exp):
    """Analyzer.calc_diff_coeff (analysis.py:460-511), loop for loop, with
    numpy's trapezoid for the removed scipy.integrate.trapz.

    Pinned by closed forms: a correlation that is constant after the phase
    shift grows linearly in time (rectangle window), and under a geometric
    window r^i the running trapezoid sum is a geometric series."""
    t_step = 2 * math.pi / en_range
    energies = np.array([0.5 * i * en_range / tnr - en_range / 2.
                         for i in range(tnr * 2)], dtype=float)
    qe = np.where((energies >= energy_limits[0]) & (energies <= energy_limits[1]))[0]
    energies = energies[qe]
    time = np.linspace(0, tnr - 1, tnr) * t_step
    diff = np.zeros((2, len(qe), tnr))
    for i in range(2):
        for j in range(len(qe)):
            en = energies.item(j)
            temp = np.zeros(tnr)
            for k in range(tnr):
                w = window_dc(k + 1, tnr)
                phi = k * t_step * en
                c_exp = math.cos(phi) - 1j * math.sin(phi)
                temp[k] = w * (c_exp * corr_dc.item(i, j, k)).real
            for k2 in range(tnr):
                diff[i, j, k2] = np.trapezoid(temp[:k2 + 1], time[:k2 + 1])
    return time, diff
