"""Seeded synthetic inputs: counter-based SplitMix64, uniform doubles from the top 53 bits.

The same stream is produced for any array length (value i depends only on (seed, i)), so the
CPU checker and the device path see bit-identical inputs. Streams are separated by seed.
"""
import numpy as np

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed, n, offset=0):
    """n 64-bit outputs of SplitMix64 at counters offset+1 .. offset+n."""
    with np.errstate(over="ignore"):
        i = np.arange(offset + 1, offset + n + 1, dtype=np.uint64)
        z = np.uint64(seed) + i * _GOLDEN
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        z = z ^ (z >> np.uint64(31))
    return z


def uniform(seed, n, lo=0.0, hi=1.0):
    """n doubles uniform in [lo, hi): lo + (hi - lo) * (top53 * 2^-53)."""
    u = (splitmix64(seed, n) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    return lo + (hi - lo) * u


def normal(seed, n):
    """Standard normals by Box-Muller on two uniform streams (seed, seed ^ 0x5bd1e995)."""
    u1 = uniform(seed, n)
    u2 = uniform(seed ^ 0x5BD1E995, n)
    u1 = np.where(u1 <= 0.0, 1.0 / 9007199254740992.0, u1)
    return np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2)


def lognormal_mean_cv(seed, n, mean, cv):
    """Lognormal with the given arithmetic mean and coefficient of variation."""
    s2 = np.log1p(cv * cv)
    mu = np.log(mean) - 0.5 * s2
    return np.exp(mu + np.sqrt(s2) * normal(seed, n))


# Stream ids used by the workloads (displacement, nonlocal strain, history, material fields).
SEED_U = 0x1001
SEED_E = 0x1002
SEED_KAPPA = 0x1003
SEED_E_MOD = 0x2001
SEED_FT = 0x2002
SEED_GF = 0x2003
SEED_PERTURB = 0x2004


def material_fields(n_cells):
    """Per-cell (E, f_t, G_f): E ~ LN(mean 30000, CV 0.1), f_t ~ LN(3, 0.1), G_f ~ LN(0.1, 0.15)."""
    return (lognormal_mean_cv(SEED_E_MOD, n_cells, 30000.0, 0.1), lognormal_mean_cv(SEED_FT, n_cells, 3.0, 0.1),
            lognormal_mean_cv(SEED_GF, n_cells, 0.1, 0.15))


def perturbed_fields(E, ft, gf, scale=0.05):
    """Seeded perturbation of the fields (the point r_obs is observed at): each value times
    exp(scale * z), z standard normal from stream SEED_PERTURB (one counter range per field)."""
    n = len(E)
    z = normal(SEED_PERTURB, 3 * n).reshape(3, n)
    return E * np.exp(scale * z[0]), ft * np.exp(scale * z[1]), gf * np.exp(scale * z[2])
