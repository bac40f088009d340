"""axbc_batch_utility_noalloc through the C ABI against axbc_batch_utility:
bit-identical per-root utilities on two calls in a row, the second with a
smaller input, on a real batch; and the size limit rejected as an error."""
import ctypes as C

import numpy as np
import pytest

from axbc_amd import Engine, Graph, _ffi

pytestmark = pytest.mark.gpu


def _utility(L, fn, e, vertex, weight, width):
    out = np.full(width, np.nan)
    v = np.ascontiguousarray(vertex, np.uint32)
    w = np.ascontiguousarray(weight, np.float64)
    rc = getattr(L, fn)(e.ctx, v.ctypes.data, w.ctypes.data, len(v), out.ctypes.data)
    assert rc == 0, L.axbc_last_error()
    return out


@pytest.mark.parametrize("width", [64, 128])
def test_noalloc_utility_matches_the_allocating_one(require_gpu, width):
    L = _ffi.hip()
    L.axbc_batch_utility_noalloc.restype = C.c_int
    L.axbc_batch_utility_noalloc.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p]
    g = Graph.kronecker(12, 16, seed=(1, 2), heuristic=1)
    deg = g.degrees()
    rng = np.random.default_rng(width)
    roots = rng.choice(np.flatnonzero(deg > 0), size=width - 3, replace=False)
    with Engine(g) as e:
        e.set_batch_width(width)
        e.batch(roots, np.full(len(roots), 1.0 / g.N))
        print(f"[noalloc] width {width}: batch of {len(roots)} roots done", flush=True)
        for count in (60000, 777, 0):
            vertex = rng.integers(0, g.N, size=count)
            weight = rng.random(count) * 3.0
            a = _utility(L, "axbc_batch_utility_noalloc", e, vertex, weight, width)
            b = _utility(L, "axbc_batch_utility", e, vertex, weight, width)
            assert np.array_equal(a, b), f"count {count}: noalloc differs from the allocating call"
            if count:
                # the value itself: sum_i weight_i * delta_k(vertex_i)
                for k in (0, len(roots) // 2, len(roots) - 1):
                    want = float(np.dot(weight, e.batch_delta(k)[vertex]))
                    assert a[k] == pytest.approx(want, rel=1e-9)
            else:
                assert (a == 0).all()
        big = np.zeros(65537, np.uint32)
        rc = L.axbc_batch_utility_noalloc(e.ctx, big.ctypes.data, np.ones(65537).ctypes.data, 65537,
                                          np.empty(width).ctypes.data)
        assert rc == -1 and b"65536" in L.axbc_last_error()
