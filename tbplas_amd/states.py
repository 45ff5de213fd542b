"""Seeded random initial states, generated on the host.

The reference draws its states inside Fortran with the compiler's intrinsic
RNG (tbplas/fortran/random.f90:14-69), which no other compiler reproduces.
Here states come from a counter-based splitmix64 stream, so the same seed
gives the same state on every machine and the states can be handed to any
implementation (CPU or GPU) unchanged.

Seeds follow the reference's per-sample rule
``seed * (i_sample + rank * n_ran_samples)`` with 1-based ``i_sample``
(tbplas/fortran/tbpm.f90:57).
"""
import numpy as np

__all__ = ["splitmix64", "uniform01", "random_state", "random_states",
           "sample_seeds"]

_GAMMA = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def _mix(z):
    z = (z ^ (z >> np.uint64(30))) * _M1
    z = (z ^ (z >> np.uint64(27))) * _M2
    return z ^ (z >> np.uint64(31))


def splitmix64(key, count, start=0):
    """Outputs start..start+count-1 of the splitmix64 stream seeded by key."""
    with np.errstate(over="ignore"):
        i = np.arange(start + 1, start + count + 1, dtype=np.uint64)
        return _mix(np.uint64(key) + i * _GAMMA)


def uniform01(key, count, start=0):
    """Uniform doubles in [0, 1) with 53 random bits."""
    return (splitmix64(key, count, start) >> np.uint64(11)).astype(np.float64) \
        * (1.0 / 9007199254740992.0)


def _key(seed):
    with np.errstate(over="ignore"):
        return int(_mix(np.uint64(int(seed) & 0xFFFFFFFFFFFFFFFF)))


def random_state(n, seed, kind="phase"):
    """One normalised complex128 state of length n.

    kind="phase": unit amplitudes with phases uniform in [0, 2 pi).
    kind="exp":   |z|^2 ~ Exp(1) and uniform phases, the distribution of the
                  reference's random_state (random.f90:56-61).
    """
    key = _key(seed)
    if kind == "phase":
        phi = 2.0 * np.pi * uniform01(key, n)
        z = np.cos(phi) + 1j * np.sin(phi)
    elif kind == "exp":
        u = uniform01(key, 2 * n)
        f, g = u[0::2], u[1::2]
        amp = np.sqrt(-np.log(1.0 - f))
        z = amp * (np.cos(2.0 * np.pi * g) + 1j * np.sin(2.0 * np.pi * g))
    else:
        raise ValueError(f"unknown state kind {kind!r}")
    return (z / np.sqrt(np.sum(np.abs(z) ** 2))).astype(np.complex128)


def sample_seeds(seed, n_samples, rank=0):
    """Per-sample seeds seed*(i_sample + rank*n_samples), i_sample = 1.."""
    return [int(seed) * (i + 1 + rank * n_samples) for i in range(n_samples)]


def random_states(n, seeds, kind="phase"):
    """States for a list of seeds, shape (len(seeds), n)."""
    out = np.empty((len(seeds), n), dtype=np.complex128)
    for i, s in enumerate(seeds):
        out[i] = random_state(n, s, kind)
    return out
