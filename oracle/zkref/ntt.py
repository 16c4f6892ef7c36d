"""
Radix-2 NTT over a prime field (oracle; test infrastructure only).

Restates src/zkgpu/ntt.py:53-203: Miller-Rabin primality with the 13 prime
bases up to 41, the smallest quadratic non-residue as the generator of every
root, omega_n = g^((m-1)/n), and the iterative decimation-in-time transform on
a bit-reversed copy, whose stage of length L uses w_L = omega_n^(n/L). The
inverse uses omega^-1 and scales by n^-1. Error messages match the
reference's so tests can check them.
"""

from __future__ import annotations

import functools

_BASES = (2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37, 41)


def is_probable_prime(n: int) -> bool:
    """Miller-Rabin over _BASES (ntt.py:53-71)."""
    if n < 2:
        return False
    for b in _BASES:
        if n % b == 0:
            return n == b
    d, s = n - 1, 0
    while d % 2 == 0:
        d //= 2
        s += 1
    for b in _BASES:
        x = pow(b, d, n)
        if x == 1 or x == n - 1:
            continue
        for _ in range(s - 1):
            x = x * x % n
            if x == n - 1:
                break
        else:
            return False
    return True


@functools.cache
def quadratic_non_residue(m: int) -> int:
    """Smallest g with Euler's criterion g^((m-1)/2) = -1 (ntt.py:75-95)."""
    if m < 3 or not is_probable_prime(m):
        raise ValueError(f"NTT modulus must be an odd prime, got {m}")
    g = 2
    while pow(g, (m - 1) // 2, m) != m - 1:
        g += 1
    return g


def get_ntt_root(n: int, m: int, non_residue: int | None = None) -> int:
    """Primitive n-th root of unity g^((m-1)/n) (ntt.py:98-124)."""
    if n < 1 or n & (n - 1):
        raise ValueError(f"NTT size must be a power of 2, got {n}")
    g0 = quadratic_non_residue(m)
    g = g0 if non_residue is None else non_residue
    if (m - 1) % n:
        raise ValueError(
            f"NTT of size {n} not supported for mod={m}: (mod-1) = {m - 1} is not divisible by n."
        )
    w = pow(g, (m - 1) // n, m)
    if n > 1 and pow(w, n // 2, m) != m - 1:
        raise ValueError(f"{g} gives no primitive {n}-th root of unity mod {m}; it must be a quadratic non-residue")
    return w


def bit_reverse_permute(a: list[int]) -> list[int]:
    """out[i] = a[rev(i)] over log2(len) bits (ntt.py:132-141)."""
    n = len(a)
    bits = n.bit_length() - 1
    out = [0] * n
    for i in range(n):
        r = 0
        x = i
        for _ in range(bits):
            r = (r << 1) | (x & 1)
            x >>= 1
        out[r] = a[i]
    return out


def ntt(a: list[int], m: int, invert: bool = False) -> list[int]:
    """The reference's transform, value for value (ntt.py:149-203)."""
    n = len(a)
    if n == 0:
        return []
    if n & (n - 1):
        raise ValueError(f"NTT length must be a power of 2, got {n}")
    w_n = get_ntt_root(n, m)
    if invert:
        w_n = pow(w_n, m - 2, m)
    if n == 1:
        return [a[0] % m]
    x = bit_reverse_permute([v % m for v in a])
    length = 2
    while length <= n:
        w_len = pow(w_n, n // length, m)
        half = length // 2
        for start in range(0, n, length):
            w = 1
            for j in range(start, start + half):
                u, v = x[j], x[j + half] * w % m
                x[j], x[j + half] = (u + v) % m, (u - v) % m
                w = w * w_len % m
        length *= 2
    if invert:
        n_inv = pow(n, m - 2, m)
        x = [v * n_inv % m for v in x]
    return x


def dft(a: list[int], m: int, invert: bool = False) -> list[int]:
    """O(n^2) definition sum_i a_i w^(ik), independent of ntt's butterfly order (for pinning)."""
    n = len(a)
    w = get_ntt_root(n, m)
    if invert:
        w = pow(w, -1, m)
    out = []
    for k in range(n):
        wk, acc, p = pow(w, k, m), 0, 1
        for v in a:
            acc = (acc + v * p) % m
            p = p * wk % m
        out.append(acc * pow(n, -1, m) % m if invert else acc)
    return out
