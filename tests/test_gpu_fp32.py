"""FP32 explicit steps (FUSS_FP32 contexts): T, fluxes, RK update and IRS in
single precision, the local dt formed in FP64, the dt minimum and
Sum (T - T^n)^2 reduced in FP64.

Checked against the FP64 path (which meets the oracle to 1e-12): after
20 steps the FP32 state agrees to the stated bound (FP32 rounding of the
state is 6e-8 relative; the bound allows for its growth over the steps),
dtglobal and the residual to the same order; the captured-graph run equals
the FP32 step loop bit for bit; state handover to FP64 consumers (get/set,
probes, wall output, multigrid transfer, defect correction) is exact.
"""
import numpy as np
import pytest

from fuss_amd import cases
from tests import cases3d as c3
from tests.test_gpu_parity import relerr

pytestmark = pytest.mark.gpu

# relative L-inf bound on T after 20 FP32 steps against FP64
FP32_RTOL = 2e-6


@pytest.fixture(scope="module")
def gpu():
    from tests.conftest import gpu_available

    if not gpu_available():
        pytest.skip("no GPU")
    from fuss_amd import gpu as g

    g.load()
    return g


FP32_CASES = {
    "cube3d_rk4_irs": lambda: cases.cube3d(24, mgl=1, n_rk=4),
    "cube3d_mg3_irs": lambda: cases.cube3d(24, mgl=3),
    "cube3d_cons": lambda: cases.cube3d(20, mgl=1, irs_beta=0.0, prim=False, vnn=0.8),
    "lattice_transient": lambda: cases.lattice(nl=2, lo=4, hi=8),
    "lattice_rk4": lambda: cases.lattice(nl=2, lo=6, hi=14, n_rk=4),
    "lattice_cons_irs_mg2": lambda: cases.lattice(nl=2, lo=4, hi=8, prim=False, irs_beta=0.5,
                                                  time_accurate=False, mgl=2),
    "plate2d_64": lambda: cases.plate2d(64),
    "plate2d_irs": lambda: cases.plate2d(40, irs_beta=0.5),
    "sweep3_rk4_irs": lambda: cases.sweep(ncase=3, lo=12, hi=28, n_rk=4, mgl=1, seed=7),
    "affine_walls": lambda: c3.affine_quadratic(
        n_rk=3, faces=[(302, cases.R(Tw=480.0)), (303, cases.R(hconv=40.0, Tref=300.0)),
                       (304, cases.R(eps=0.8, Tref=300.0)), (305, cases.R(eps=0.6, hconv=20.0, Tref=320.0)),
                       (301, cases.R(qw=2e3)), (300, cases.R())]),
    "cube3d_tridiag": lambda: _tridiag(cases.cube3d(24, mgl=1, n_rk=4)),
}


def _tridiag(c):
    c.params.irs_mode = 1
    return c


def run_steps(gpu, c, precision, n=20, level=None):
    s = gpu.Solver(c, precision=precision)
    lv = level or c.params.mgl
    rs = [s.explicit_step(lv) for _ in range(n)]
    T = [s.get("T", lv, b + 1) for b in range(c.nb)]
    clock = s.clock(lv)
    s.close()
    return T, rs, clock


@pytest.mark.parametrize("name", list(FP32_CASES))
def test_fp32_steps_vs_fp64(gpu, name):
    c = FP32_CASES[name]()
    T64, r64, k64 = run_steps(gpu, c, gpu.FUSS_FP64)
    T32, r32, k32 = run_steps(gpu, FP32_CASES[name](), gpu.FUSS_FP32)
    err = max(relerr(a, b) for a, b in zip(T32, T64))
    dterr = max(abs(a.dtglobal - b.dtglobal) / b.dtglobal for a, b in zip(r32, r64))
    res = np.array([np.sqrt(r.res_sq) for r in r64])
    reserr = max(abs(np.sqrt(a.res_sq) - np.sqrt(b.res_sq)) for a, b in zip(r32, r64)) / res.max()
    print(f"{name}: FP32 vs FP64 after 20 steps: T {err:.2e}, dtglobal {dterr:.2e}, residual {reserr:.2e}",
          flush=True)
    assert err < FP32_RTOL
    assert dterr < 1e-5
    assert reserr < 1e-2
    assert k32[1] == k64[1] and abs(k32[0] - k64[0]) <= 1e-5 * max(abs(k64[0]), 1.0)
    # not silently FP64: the single-precision rounding shows
    assert err > 1e-10


def test_fp32_graph_run_equals_step_loop(gpu):
    """fuss_run on captured graphs (60 steps) equals 60 FP32 fuss_step calls."""
    c = cases.lattice(nl=2, lo=4, hi=8, n_rk=4)
    s = gpu.Solver(c, precision=gpu.FUSS_FP32)
    k, hr, ht, r = s.run_level(1, 60, 60, -1.0, 1e300)
    Tg = [s.get("T", 1, b + 1) for b in range(c.nb)]
    s.close()
    s = gpu.Solver(cases.lattice(nl=2, lo=4, hi=8, n_rk=4), precision=gpu.FUSS_FP32)
    s.graphs(False)
    k2, hr2, ht2, r2 = s.run_level(1, 60, 60, -1.0, 1e300)
    Tl = [s.get("T", 1, b + 1) for b in range(c.nb)]
    s.close()
    assert k == k2 == 60
    for a, b in zip(Tg, Tl):
        assert np.array_equal(a, b)
    assert np.array_equal(hr, hr2) and np.array_equal(ht, ht2)


def test_fp32_state_handover(gpu):
    """get_T / set_T / probes / wall output read the FP32 state; set_T and
    multigrid transfer write it; continuing FP32 steps after a get/set round
    trip equals not reading at all (the conversions are exact)."""
    mk = lambda mgl: cases.cube3d(24, mgl=mgl, n_rk=3)  # noqa: E731
    a = gpu.Solver(mk(1), precision=gpu.FUSS_FP32)
    b = gpu.Solver(mk(1), precision=gpu.FUSS_FP32)
    for _ in range(5):
        a.explicit_step(1)
        b.explicit_step(1)
    T = b.get("T", 1)
    assert float(b.probe([(1, 3, 4, 5)])[0]) == T[4, 3, 2]
    assert np.array_equal(T.astype(np.float32).astype(np.float64), T)  # an FP32 state
    b.set_T(T, 1)
    b.wall(1)
    for _ in range(5):
        a.explicit_step(1)
        b.explicit_step(1)
    assert np.array_equal(a.get("T", 1), b.get("T", 1))
    for s in (a, b):
        s.close()
    # coarse FP32 steps, prolongation (an FP64 transfer), fine FP32 steps
    a = gpu.Solver(mk(2), precision=gpu.FUSS_FP32)
    ref = gpu.Solver(mk(2), precision=gpu.FUSS_FP64)
    for _ in range(10):
        a.explicit_step(2)
    ref.set_T(a.get("T", 2), 2)
    a.prolong(1)
    ref.prolong(1)
    assert np.array_equal(a.get("T", 1), ref.get("T", 1))
    for _ in range(3):
        a.explicit_step(1)
        ref.explicit_step(1)
    assert relerr(a.get("T", 1), ref.get("T", 1)) < FP32_RTOL
    for s in (a, ref):
        s.close()


def test_fp32_context_refine_monitor_fp64(gpu):
    """Defect correction in an FP32 context: the monitor steps and the
    defect are FP64, so the result meets the reference's residual test
    exactly as in an FP64 context."""
    from tests.test_refine import refined

    T32c, r32c, _ = refined(gpu, cases.cube3d(n=16, mgl=1, n_rk=3), gpu.FUSS_FP32, ctx_precision=gpu.FUSS_FP32)
    T64c, r64c, _ = refined(gpu, cases.cube3d(n=16, mgl=1, n_rk=3), gpu.FUSS_FP32, ctx_precision=gpu.FUSS_FP64)
    assert r32c.converged and r32c.res <= 1e-10
    assert relerr(T32c[0], T64c[0]) < 1e-14


def test_fp32_unsupported_paths(gpu):
    """Hooks are FP64 only: an FP32 context refuses them with
    FUSS_EUNSUPPORTED instead of running something else."""
    s = gpu.Solver(cases.plate2d(16), precision=gpu.FUSS_FP32)
    with pytest.raises(gpu.FussError) as ei:
        s.explicit_step(1, hook=lambda *a: None)
    assert ei.value.code == gpu.FUSS_EUNSUPPORTED
    s.explicit_step(1)  # the context still steps
    s.close()
