"""Autotuner for the launch shape of the propagation kernels.

For a prepared Hamiltonian, a state count and a precision it times a few
Chebyshev steps (HIP events on the library stream, after warm-up steps) for
every candidate (persistent blocks per CU, LDS staging of the slice rows)
and keeps the fastest; for a Hamiltonian with spatial tiles it then times
the best shape with three, four and five terms per pass and keeps the
fastest.  A shape is (blocks per CU, staging, terms per pass: 1, 3, 4 or
5).
Results are cached per (device name, library build, instances, rows, stored
elements, widest slice, value layout, row order and tile count, states,
precision) in memory and in a JSON file
(``$TBPLAS_AMD_TUNE_CACHE``, default ``~/.cache/tbplas_amd/tune.json``).
"""
import json
import os
import threading

import numpy as np

from . import _lib

__all__ = ["autotune", "apply", "CANDIDATES"]

CANDIDATES = [(bpc, stage) for bpc in (4, 8, 12, 16) for stage in (True, False)]
# untuned: blocks per CU and staging; the terms per pass are those the
# tiles of auto_tile_nm are sized for (sample.pass_terms)
DEFAULT = (16, True)

_lock = threading.Lock()
_mem = {}


def _cache_path():
    return os.environ.get("TBPLAS_AMD_TUNE_CACHE",
                          os.path.join(os.path.expanduser("~"), ".cache", "tbplas_amd",
                                       "tune.json"))


def _load_disk():
    try:
        with open(_cache_path()) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _save_disk(d):
    p = _cache_path()
    try:
        os.makedirs(os.path.dirname(p), exist_ok=True)
        tmp = p + ".tmp"
        with open(tmp, "w") as f:
            json.dump(d, f, indent=1, sort_keys=True)
        os.replace(tmp, p)
    except OSError:
        pass


def _key(hams, K):
    import torch
    infos = [H.info() for H in hams]
    H0 = hams[0]
    dev = torch.cuda.get_device_name(H0.handle.device)
    st = os.stat(_lib.LIB_PATH)          # a rebuilt library is tuned afresh
    lib = f"{st.st_size}.{int(st.st_mtime)}"
    n = sum(i["n"] for i in infos)
    pad = sum(i["padded"] for i in infos)
    w = max(i["max_width"] for i in infos)
    real = all(i["real_values"] for i in infos)
    # the row order and the tiles change which shape wins
    order = "|".join(f"{int(getattr(H, 'ordered', False))}.{getattr(H, 'ntiles', 0)}"
                     for H in hams)
    return (f"{dev}|lib={lib}|inst={len(hams)}|n={n}|pad={pad}|w={w}|real={int(real)}"
            f"|order={order}|K={K}|p={H0.precision}")


def autotune(H, K, bes, steps=6, warmup=2, candidates=None, use_cache=True):
    """Pick (blocks_per_cu, stage, three terms per pass) for Hamiltonian H
    (or a list of Hamiltonians run as one batch) and K states each; returns
    (best, timings) and leaves the handle set to the best shape."""
    from .tbpm import DosRun
    hams = list(H) if isinstance(H, (list, tuple)) else [H]
    H = hams[0]
    key = _key(hams, K)
    with _lock:
        if use_cache:
            if key in _mem:
                best = tuple(_mem[key])
                _set(H.handle, best)
                return best, None
            disk = _load_disk()
            if key in disk and len(disk[key]) == 3:
                best = (int(disk[key][0]), bool(disk[key][1]), int(disk[key][2]))
                _mem[key] = best
                _set(H.handle, best)
                return best, None
        rng = np.random.default_rng(0)
        psis = []
        for Hi in hams:
            psi = np.exp(2j * np.pi * rng.random((K, Hi.n)))
            psis.append(psi / np.linalg.norm(psi, axis=1, keepdims=True))
        def timed(shape):
            _set(H.handle, shape)
            run = DosRun(hams, bes, psis, warmup + steps)
            run.advance(warmup)
            ms = run.advance(steps) / steps
            run.close()
            return ms

        timings = {}
        for cand in candidates or CANDIDATES:
            timings[cand + (1,)] = timed(cand + (1,))
        best = min(timings, key=timings.get)
        if all(getattr(Hi, "tiled", False) for Hi in hams):
            for terms in (3, 4, 5):
                blocked = best[:2] + (terms,)
                timings[blocked] = timed(blocked)
            best = min(timings, key=timings.get)
        _set(H.handle, best)
        _mem[key] = best
        if use_cache:
            disk = _load_disk()
            disk[key] = [best[0], best[1], best[2]]
            _save_disk(disk)
        return best, timings


def _set(handle, shape):
    handle.set_tuning(shape[0], shape[1])
    handle.set_blocking(shape[2] > 1, terms=max(int(shape[2]), 3))


def apply(H, K, bes, enabled=True):
    """Tune when enabled, otherwise use the default shape (H: one
    Hamiltonian or the list of a batch)."""
    if enabled:
        return autotune(H, K, bes)[0]
    from .sample import pass_terms
    H0 = H[0] if isinstance(H, (list, tuple)) else H
    prec = "fp32" if H0.precision == _lib.TBG_FP32 else "fp64"
    shape = DEFAULT + (pass_terms(K, prec),)
    _set(H0.handle, shape)
    return shape
