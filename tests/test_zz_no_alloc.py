"""fuss_wall_noalloc: the wall output staged in the library's static per-GPU
array (no allocation) against fuss_wall, which allocates its staging.

Two calls in a row on different contexts, the second with fewer BC
entries (the array keeps the first call's contents beyond the second's
range), each equal to fuss_wall bit for bit; a level larger than the
array is refused with FUSS_EARG before any device work."""
import ctypes as C

import numpy as np
import pytest

from fuss_amd import cases
from tests import cases3d as c3

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gpu():
    from tests.conftest import gpu_available

    if not gpu_available():
        pytest.skip("no GPU")
    from fuss_amd import gpu as g

    g.load()
    return g


def wall_noalloc(gpu, s, level=1):
    L = gpu.load()
    n = int(s.case.bcs[level - 1].n)
    Tw, qw = np.zeros(n), np.zeros(n)
    code = L.fuss_wall_noalloc(s.h, level, Tw.ctypes.data_as(C.c_void_p), qw.ctypes.data_as(C.c_void_p))
    return code, Tw, qw


def box_case():
    return c3.box(n=(14, 10, 8), faces=[(302, cases.R(Tw=600.0)), (305, cases.R(eps=0.7, hconv=35.0, Tref=300.0)),
                                        (303, cases.R(hconv=20.0, Tref=300.0)), (304, cases.R(eps=0.8, Tref=300.0)),
                                        (301, cases.R(qw=-1e4)), (0, cases.R())], T0=520.0)


def test_wall_noalloc_matches_wall(gpu):
    L = gpu.load()
    L.fuss_wall_noalloc.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]
    results = []
    for mk in (box_case, lambda: cases.plate2d(12)):  # second call: fewer entries
        c = mk()
        s = gpu.Solver(c)
        for _ in range(5):
            s.explicit_step(1)
        Tw0, qw0 = s.wall(1)
        code, Tw, qw = wall_noalloc(gpu, s)
        assert code == 0
        assert np.array_equal(Tw, Tw0) and np.array_equal(qw, qw0)
        results.append(Tw.size)
        s.close()
    assert results[1] < results[0]


def test_wall_noalloc_rejects_oversize(gpu):
    L = gpu.load()
    L.fuss_wall_noalloc.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]
    c = cases.plate2d(400)  # 2 x 400 x 400 + 4 x 400 entries > FUSS_WALL_NOALLOC_MAX
    assert c.bcs[0].n > 262144
    s = gpu.Solver(c)
    code, _, _ = wall_noalloc(gpu, s)
    assert code == gpu.FUSS_EARG
    # the context is untouched: steps and the allocating wall output still work
    s.explicit_step(1)
    Tw, _ = s.wall(1)
    assert Tw.size == c.bcs[0].n
    s.close()
