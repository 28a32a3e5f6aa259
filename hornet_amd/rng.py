"""Host-side counter-based streams (same definition as the device generators in
hornet_amd/csrc/gen.hip): draw(key, c) = SplitMix64 finaliser of
key + (c+1)*golden, key = finaliser(seed + golden)."""
from __future__ import annotations

import numpy as np

GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def mix64(z: np.ndarray) -> np.ndarray:
    z = np.asarray(z, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
    return z ^ (z >> np.uint64(31))


def stream_key(seed: int) -> np.uint64:
    with np.errstate(over="ignore"):
        return mix64(np.uint64(seed) + GOLDEN)


def draws(key, first: int, count: int) -> np.ndarray:
    c = np.arange(first + 1, first + count + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        return mix64(np.uint64(key) + c * GOLDEN)


def uniform_below(r: np.ndarray, n: int) -> np.ndarray:
    return ((r >> np.uint64(32)) * np.uint64(n)) >> np.uint64(32)


def pick_sources(degrees: np.ndarray, k: int, seed: int) -> np.ndarray:
    """k vertices with degree > 0: draws j = 0,1,...; v = uniform_below(draw(j), n)."""
    n = degrees.size
    key = stream_key(seed)
    out: list[int] = []
    j = 0
    while len(out) < k:
        if j > 64 * n + 64:
            raise ValueError("graph has no non-isolated vertex")
        v = uniform_below(draws(key, j, 1024), n).astype(np.int64)
        for x in v:
            if degrees[x] > 0:
                out.append(int(x))
                if len(out) == k:
                    break
        j += 1024
    return np.array(out, dtype=np.int32)
