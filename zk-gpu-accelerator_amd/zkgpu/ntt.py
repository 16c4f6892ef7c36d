"""
Roots of unity for power-of-two NTTs, and the checks that precede every transform.

Host logic behind Field256.ntt: the same derivation and the same error
messages as the reference (src/zkgpu/ntt.py:53-124): omega_n = g^((m-1)/n)
with g the smallest quadratic non-residue (5 for BN254's r), found with a
deterministic Miller-Rabin test. The transforms themselves run on the GPU.
"""

from __future__ import annotations

import functools

_WITNESSES = (2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37, 41)


def _is_probable_prime(n: int) -> bool:
    if n < 2:
        return False
    small = next((w for w in _WITNESSES if n % w == 0), None)
    if small is not None:
        return n == small
    odd, twos = n - 1, 0
    while not odd & 1:
        odd >>= 1
        twos += 1
    for w in _WITNESSES:
        y = pow(w, odd, n)
        if y in (1, n - 1):
            continue
        for _ in range(twos - 1):
            y = y * y % n
            if y == n - 1:
                break
        else:
            return False
    return True


@functools.cache
def quadratic_non_residue(mod: int) -> int:
    """Smallest g whose Legendre symbol is -1; ValueError unless mod is an odd prime."""
    if mod < 3 or not _is_probable_prime(mod):
        raise ValueError(f"NTT modulus must be an odd prime, got {mod}")
    half = (mod - 1) // 2
    return next(g for g in range(2, mod) if pow(g, half, mod) == mod - 1)


def get_ntt_root(n: int, mod: int, non_residue: int | None = None) -> int:
    """A primitive n-th root of unity mod *mod* for a power-of-two n."""
    if n < 1 or n & (n - 1):
        raise ValueError(f"NTT size must be a power of 2, got {n}")
    default = quadratic_non_residue(mod)
    g = default if non_residue is None else non_residue
    if (mod - 1) % n:
        raise ValueError(
            f"NTT of size {n} not supported for mod={mod}: (mod-1) = {mod - 1} is not divisible by n."
        )
    omega = pow(g, (mod - 1) // n, mod)
    if n > 1 and pow(omega, n // 2, mod) != mod - 1:
        raise ValueError(f"{g} gives no primitive {n}-th root of unity mod {mod}; it must be a quadratic non-residue")
    return omega
