"""Synthetic inputs of the five benchmark configurations (seeded, reproducible).

CN(0, s^2) means real and imaginary parts each N(0, s^2 / 2).
  1. one 20x20 complex A = (B + B^T)/2, B = N(0,1) + i N(0,1), plain hafnian (seed 137)
  2. 4096 real symmetric 24x24, U[-1, 1] symmetrised, plus 16 all-ones 24x24 (seed 138)
  3. one 40x40 complex loop hafnian: upper triangle CN(0, 1/40) mirrored, diagonal CN(0, 1)
     (seed 139)
  4. 1024 hafnian_repeated requests: 8x8, upper triangle CN(0, 1/8) mirrored,
     rpt ~ Multinomial(32, uniform), mu ~ CN(0, 1), loop=True (seed 140)
  5. the serving stream: even n uniform in [8, 40], loop with probability 1/2,
     A = (B + B^T)/2 with B ~ CN(0, 1/n), Poisson arrivals (seed 141)
"""
import numpy as np


def _cn(rng, scale2, size):
    s = np.sqrt(scale2 / 2.0)
    return rng.normal(scale=s, size=size) + 1j * rng.normal(scale=s, size=size)


def _sym_upper(rng, n, scale2):
    U = _cn(rng, scale2, (n, n))
    A = np.triu(U, 1)
    return A + A.T


def config1(seed=137):
    rng = np.random.default_rng(seed)
    B = rng.normal(size=(20, 20)) + 1j * rng.normal(size=(20, 20))
    return (B + B.T) / 2


def config2(count=4096, ones=16, seed=138):
    rng = np.random.default_rng(seed)
    M = rng.uniform(-1, 1, size=(count, 24, 24))
    mats = [(m + m.T) / 2 for m in M]
    return mats + [np.ones((24, 24)) for _ in range(ones)]


def config3(n=40, seed=139):
    rng = np.random.default_rng(seed)
    A = _sym_upper(rng, n, 1.0 / n)
    A[np.diag_indices(n)] = _cn(rng, 1.0, n)
    return A


def config4(count=1024, seed=140):
    """List of (A, rpt, mu)."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        A = _sym_upper(rng, 8, 1.0 / 8)
        rpt = rng.multinomial(32, [1.0 / 8] * 8)
        mu = _cn(rng, 1.0, 8)
        out.append((A, [int(r) for r in rpt], mu))
    return out


def config5(count=10000, seed=141, nmin=8, nmax=40):
    """List of (A, loop) in arrival order."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        n = 2 * int(rng.integers(nmin // 2, nmax // 2 + 1))
        loop = bool(rng.random() < 0.5)
        B = _cn(rng, 1.0 / n, (n, n))
        out.append(((B + B.T) / 2, loop))
    return out


def arrivals(count, qps, seed=141):
    """Poisson arrival times (s) for ``count`` requests at rate ``qps``."""
    rng = np.random.default_rng(seed + 1_000_003)
    return np.cumsum(rng.exponential(1.0 / qps, size=count))


def w_glynn(n, real=False):
    """Products-only FP64 flops of the reference algorithm for one even-n request with every
    repetition 1: 2^(h-1) terms, each h-1 products of n x n (8 flops per complex
    multiply-add, 2 per real one)."""
    h = n // 2
    return 2 ** (h - 1) * (2.0 if real else 8.0) * n ** 3 * max(h - 1, 0)
