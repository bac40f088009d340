"""The engine's Kronecker generator against fixtures written by the
reference Graph500 generator (tests/golden/make_generator_fixtures.py), plus
mathematical pins of its MRG skip-ahead."""
import ctypes as C
import json
import os

import numpy as np
import pytest

from axbc_amd import kronecker_edge_list
from axbc_amd import _ffi
from tests.golden.make_generator_fixtures import digest

FIX = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "generator.json")))


@pytest.mark.parametrize("case", FIX["cases"], ids=lambda c: f"s{c['scale']}-m{c['nedges']}-{c['seed1']}")
def test_matches_reference_generator(case):
    e = kronecker_edge_list(case["scale"], case["nedges"], case["seed1"], case["seed2"])
    d = digest(e)
    for k in ("count", "sha256", "head", "tail", "self_loops"):
        assert d[k] == case[k], k


def test_thread_count_does_not_change_output():
    a = kronecker_edge_list(12, 65536, 5, 6, threads=1)
    b = kronecker_edge_list(12, 65536, 5, 6, threads=7)
    assert np.array_equal(np.sort(a, axis=0), np.sort(b, axis=0))
    assert digest(a) == digest(b)


class Mrg(C.Structure):
    _fields_ = [("z", C.c_uint32 * 5)]


def test_mrg_skip_equals_stepping_and_composes():
    L = _ffi.host()
    a, b = Mrg(), Mrg()
    L.axbc_mrg_seed_from_user(11, 22, C.byref(a))
    C.memmove(C.byref(b), C.byref(a), C.sizeof(Mrg))
    for _ in range(1000):
        L.axbc_mrg_next_uint(C.byref(a))
    L.axbc_mrg_skip(C.byref(b), 0, 0, 1000)
    assert list(a.z) == list(b.z)
    # skip(x) then skip(y) == skip(x + y) across the 64-bit word boundary
    c, d = Mrg(), Mrg()
    L.axbc_mrg_seed_from_user(3, 4, C.byref(c))
    C.memmove(C.byref(d), C.byref(c), C.sizeof(Mrg))
    L.axbc_mrg_skip(C.byref(c), 0, 5, (1 << 64) - 7)
    L.axbc_mrg_skip(C.byref(c), 0, 0, 9)
    L.axbc_mrg_skip(C.byref(d), 0, 6, 2)
    assert list(c.z) == list(d.z)


def test_mrg_recurrence():
    # z_new = (107374182 z1 + 104480 z5) mod (2^31 - 1)
    L = _ffi.host()
    s = Mrg()
    L.axbc_mrg_seed_from_user(1, 2, C.byref(s))
    z = list(s.z)
    for _ in range(200):
        nxt = (107374182 * z[0] + 104480 * z[4]) % (2 ** 31 - 1)
        assert L.axbc_mrg_next_uint(C.byref(s)) == nxt
        z = [nxt] + z[:4]


def test_binomial_moments():
    # Unbiased mean.  The variance is about 0.75-1.0 of n p q: the reference
    # generator's BTRD step 1 divides by (0.5 + |u|) where Hoermann's paper
    # has (0.5 - |u|), and the graphs must reproduce that.
    L = _ffi.host()
    s = Mrg()
    L.axbc_mrg_seed_from_user(9, 9, C.byref(s))
    for n, p in [(1000, 0.3), (100000, 0.57), (50, 0.05), (2000, 0.004)]:
        x = np.array([L.axbc_btrd_binomial(n, p, C.byref(s)) for _ in range(4000)], np.float64)
        assert (x >= 0).all() and (x <= n).all()
        mu, var = n * p, n * p * (1 - p)
        assert abs(x.mean() - mu) < 5 * np.sqrt(var / x.size) + 1e-9
        assert 0.65 < x.var() / var < 1.15
