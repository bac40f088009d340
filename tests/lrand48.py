"""POSIX drand48 family in pure Python (test helper).

X_{n+1} = (0x5DEECE66D * X_n + 0xB) mod 2^48; srand48(seed) sets
X = (seed << 16) | 0x330E; lrand48() returns X >> 17.  Independent of the
C library; test_oracle.py checks it against libc's lrand48.
"""
from __future__ import annotations


class Lrand48:
    def __init__(self, seed: int = 4343543):
        self.x = ((seed & 0xFFFFFFFF) << 16) | 0x330E

    def next(self) -> int:
        self.x = (0x5DEECE66D * self.x + 0xB) % (1 << 48)
        return self.x >> 17

    def uniform_int(self, hi: int, lo: int = 0) -> int:
        """lrand48() % (hi - lo + 1) + lo in uint32, as the reference's
        uniform_int (src/axbc2d.c:2156-2164); an empty modulus (hi - lo + 1
        wrapping to 0) returns the raw draw + lo."""
        m = (hi - lo + 1) & 0xFFFFFFFF
        r = self.next()
        return ((r % m) + lo) & 0xFFFFFFFF if m else (r + lo) & 0xFFFFFFFF
