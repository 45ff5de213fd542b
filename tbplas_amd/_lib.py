"""ctypes binding of the C ABI in ``include/tbplas_amd.h``.

Device buffers are PyTorch tensors on the GPU, used only as storage: their
``data_ptr()`` is what crosses the boundary.  There is no CPU fallback: if
the HIP library is missing or no GPU is visible the calls raise.
"""
import ctypes
import os
import threading

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libtbplas_amd.so")

TBG_OK, TBG_ENORM, TBG_EINVAL, TBG_EOOM, TBG_EHIP = 0, 1, 2, 3, 4
TBG_FP64, TBG_FP32 = 0, 1

_P = ctypes.c_void_p
_I64 = ctypes.c_int64
_I = ctypes.c_int
_D = ctypes.c_double

# name -> argtypes (all return int status unless listed in _RESTYPE)
_SIGS = {
    "tbg_version": [],
    "tbg_last_error": [],
    "tbg_create": [ctypes.POINTER(_P), _I],
    "tbg_destroy": [_P],
    "tbg_set_graph": [_P, _I],
    "tbg_set_tuning": [_P, _I, _I],
    "tbg_set_real_layout": [_P, _I],
    "tbg_ham_is_real": [_P, ctypes.POINTER(_I)],
    "tbg_ham_dxy_count": [_P, _I64, _I64, _P, _P, _P, _D, _P, ctypes.POINTER(_I64)],
    "tbg_ham_dxy_fill": [_P, _I64, _I64, _P, _P, _P, _P, _P, _D, _P, _P, _P, _P, _P],
    "tbg_ham_csr_count": [_P, _I64, _I64, _P, _P, _P, _P, _P, ctypes.POINTER(_I64)],
    "tbg_ham_csr_fill": [_P, _I64, _I64, _P, _P, _P, _P, _P, _P, _P],
    "tbg_get_rescale": [_P, _I64, _I64, _P, _P, _P, _P, ctypes.POINTER(_D)],
    "tbg_ham_create": [_P, _I64, _P, _P, _P, _P, _P, _D, _I, ctypes.POINTER(_P)],
    "tbg_ham_create_ordered": [_P, _I64, _P, _P, _P, _P, _P, _D, _I, _P, _P, _I64,
                               ctypes.POINTER(_P)],
    "tbg_set_blocking": [_P, _I, _I],
    "tbg_ham_blocking_info": [_P, _P, _I, _I, ctypes.POINTER(_I), ctypes.POINTER(_D)],
    "tbg_ham_blocking_launch": [_P, _P, _I, _I, ctypes.POINTER(_I), ctypes.POINTER(_I)],
    "tbg_ham_blocking_bytes": [_P, _P, _I, _I, _P],
    "tbg_set_kernel_timing": [_P, _I],
    "tbg_kernel_times": [_P, _I, _P, _P, _P, ctypes.POINTER(_I)],
    "tbg_ham_destroy": [_P],
    "tbg_ham_info": [_P, ctypes.POINTER(_I64), ctypes.POINTER(_I64),
                     ctypes.POINTER(_I64), ctypes.POINTER(_I)],
    "tbg_corr_dos": [_P, _P, _P, _I, _P, _I, _I, _I, _D, _P, ctypes.POINTER(_I)],
    "tbg_corr_ldos": [_P, _P, _P, _I, _P, _I, _I, _I, _D, _P, ctypes.POINTER(_I)],
    "tbg_corr_dos_noalloc": [_P, _P, _P, _I, _P, _I, _I, _I, _D, _P, ctypes.POINTER(_I)],
    "tbg_psi_t": [_P, _P, _P, _I, _P, _P, _I, _I, _I, _D, _P, ctypes.POINTER(_I)],
    "tbg_quasi_eigenstates": [_P, _P, _P, _I, _P, _I, _I, _D, _P, _I, _I, _D, _P,
                              ctypes.POINTER(_I)],
    "tbg_hall_mu": [_P, _P, _P, _I, _I, _I, _P],
    "tbg_haydock_coef": [_P, _P, _P, _I, _I, _D, _P, _P],
    "tbg_supercell_hops": [_P, _P, _P, _I, _P, _P, _I, _P, _I64, _P, _I64, _P, _P, _I64, _P, _P,
                           _P, _P, ctypes.POINTER(_I64)],
    "tbg_corr_dyn_pol": [_P, _P, _P, _I, _P, _I, _P, _I, _P, _I, _I, _P, _I, _P, _I, _D,
                         _P, ctypes.POINTER(_I)],
    "tbg_corr_ac": [_P, _P, _P, _I, _P, _I, _P, _I, _P, _I, _I, _I, _D, _P,
                    ctypes.POINTER(_I)],
    "tbg_corr_dc": [_P, _P, _P, _I, _P, _I, _I, _D, _P, _P, _I, _I, _D, _P, _P,
                    ctypes.POINTER(_I)],
    "tbg_cheb_step": [_P, _P, _P, _I, _P, _I, _I],
    "tbg_cheb_poly": [_P, _P, _P, _I, _P, _I],
    "tbg_corr_dos_batch": [_P, _I, _P, _P, _I, _P, _I, _I, _I, _D, _P,
                           ctypes.POINTER(_I), ctypes.POINTER(_I)],
    "tbg_peierls_host": [_I64, _P, _P, _P, _P, _P, _D, _I],
    "tbg_graphene_hops": [_I64, _I64, _D, _I, _I, _D, _P, _I64, _P, _P, _P, _P, _P,
                          ctypes.POINTER(_I64), ctypes.POINTER(_I64)],
    "tbg_dos_begin": [_P, _I, _P, _P, _I, _P, _I, _I, ctypes.POINTER(_P)],
    "tbg_run_advance": [_P, _I, ctypes.POINTER(_D)],
    "tbg_dos_end": [_P, _I, _D, _P, ctypes.POINTER(_I), ctypes.POINTER(_I)],
    "tbg_run_destroy": [_P],
    "tbg_run_blocking_bytes": [_P, _P, ctypes.POINTER(_I)],
    "tbg_mlp_create": [_P, _I, _P, ctypes.POINTER(_P)],
    "tbg_mlp_destroy": [_P],
    "tbg_mlp_num_params": [_P, ctypes.POINTER(_I64)],
    "tbg_mlp_set_params": [_P, _P, _I64],
    "tbg_mlp_get_params": [_P, _P, _I64],
    "tbg_mlp_forward": [_P, _P, _I, _P],
    "tbg_mlp_gradients": [_P, _P, _P, _I, _P, ctypes.POINTER(_D)],
    "tbg_mlp_train": [_P, _P, _P, _I, _I, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                      ctypes.c_float, ctypes.POINTER(_D)],
    "tbg_gemm_bf16": [_P, _I, _I, _I, _P, _P, _P],
}
_RESTYPE = {"tbg_last_error": ctypes.c_char_p}

_lock = threading.Lock()
_lib = None


class TBGError(RuntimeError):
    """A library call failed (status code in ``.code``)."""

    def __init__(self, code, msg):
        super().__init__(msg)
        self.code = code


class NormError(TBGError):
    """The wave-function norm check failed (the reference STOPs here,
    tbplas/fortran/funcs.f90:175-192)."""

    def __init__(self, code, msg, step):
        super().__init__(code, msg)
        self.step = step


def lib():
    """Load the HIP library (raises if it has not been built)."""
    global _lib
    with _lock:
        if _lib is None:
            # torch's wheel bundles its own HIP runtime (file libamdhip64.so,
            # soname libamdhip64.so.7).  Load it first so this library binds
            # to the same runtime; loaded the other way round, torch would map
            # a second runtime and one of the two sees no device.
            import torch  # noqa: F401
            if not os.path.exists(LIB_PATH):
                raise ImportError(
                    f"{LIB_PATH} not found: build the HIP extension first "
                    "(python -m tbplas_amd.build)")
            L = ctypes.CDLL(LIB_PATH)
            for name, args in _SIGS.items():
                fn = getattr(L, name)
                fn.argtypes = args
                fn.restype = _RESTYPE.get(name, ctypes.c_int)
            _lib = L
    return _lib


def exported_symbols():
    return list(_SIGS)


def check(rc, fail_step=None):
    if rc == TBG_OK:
        return
    msg = lib().tbg_last_error().decode(errors="replace")
    if rc == TBG_ENORM:
        raise NormError(rc, msg, None if fail_step is None else fail_step.value)
    raise TBGError(rc, msg)


def torch_cuda():
    """Return torch after checking a GPU is visible (no CPU fallback)."""
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("tbplas_amd needs a ROCm GPU: torch.cuda.is_available() is False")
    return torch


class Handle:
    """One library handle (device + stream).  Calls are synchronous."""

    def __init__(self, device=0, use_graph=True):
        torch_cuda()
        h = _P()
        check(lib().tbg_create(ctypes.byref(h), int(device)))
        self._h = h
        self.device = int(device)
        self.set_graph(use_graph)

    @property
    def ptr(self):
        return self._h

    def set_graph(self, enable):
        check(lib().tbg_set_graph(self._h, 1 if enable else 0))

    def set_tuning(self, blocks_per_cu=16, stage=True):
        check(lib().tbg_set_tuning(self._h, int(blocks_per_cu), 1 if stage else 0))
        self.tuning = (int(blocks_per_cu), bool(stage))

    def set_blocking(self, enable=True, max_rows=0, terms=3):
        """Several Chebyshev terms per pass (terms = 3, 4 or 5) for
        Hamiltonians with spatial tiles (same states; off runs one term per
        pass); max_rows caps the rows one workgroup owns (0: sized from the
        LDS)."""
        if terms not in (3, 4, 5):
            raise ValueError("terms per pass must be 3, 4 or 5")
        check(lib().tbg_set_blocking(self._h, int(terms) if enable else 0, int(max_rows)))

    def set_kernel_timing(self, enable=True):
        """Bracket every propagation and correlation launch by HIP events
        (graph replay is skipped while on); read with kernel_times()."""
        check(lib().tbg_set_kernel_timing(self._h, 1 if enable else 0))

    def kernel_times(self):
        """{kind: (total ms, launches)} of the launches timed since the last
        call (synchronises the handle's stream).  Kinds: "term" (one-term
        kernel), "ac_corr", "dc_qe", "dc_dot", "gram" (the Hall moments'
        Gram block), and "pass<mode>x<terms>" for
        the tile pass (mode 0 continue, 1 first terms from v_0, 2 last
        terms, 3 with the norm, 4 with dot and norm)."""
        cap = 64
        kinds = np.zeros(cap, dtype=np.int32)
        ms = np.zeros(cap, dtype=np.float64)
        cnt = np.zeros(cap, dtype=np.int64)
        nk = ctypes.c_int()
        check(lib().tbg_kernel_times(self._h, cap, kinds.ctypes.data, ms.ctypes.data,
                                     cnt.ctypes.data, ctypes.byref(nk)))
        names = {1: "term", 2: "ac_corr", 3: "dc_qe", 4: "dc_dot", 5: "gram"}
        out = {}
        for i in range(min(nk.value, cap)):
            k = int(kinds[i])
            name = names.get(k) or f"pass{(k - 100) // 10}x{(k - 100) % 10}"
            out[name] = (float(ms[i]), int(cnt[i]))
        return out

    def set_real_layout(self, enable=True):
        """Store real values for Hamiltonians without imaginary parts
        (applies to Hamiltonians created afterwards)."""
        check(lib().tbg_set_real_layout(self._h, 1 if enable else 0))

    def close(self):
        if getattr(self, "_h", None):
            lib().tbg_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_handles = {}


def default_handle(device=0):
    h = _handles.get(device)
    if h is None:
        h = Handle(device)
        _handles[device] = h
    return h


def dev(a, dtype=None, device=0):
    """Host array -> contiguous device tensor (storage only)."""
    torch = torch_cuda()
    a = np.ascontiguousarray(a if dtype is None else np.asarray(a, dtype=dtype))
    return torch.from_numpy(a).to(f"cuda:{device}")


def ptr(t):
    return ctypes.c_void_p(t.data_ptr() if t is not None and t.numel() > 0 else 0)


def sync(device=0):
    torch_cuda().cuda.synchronize(device)
