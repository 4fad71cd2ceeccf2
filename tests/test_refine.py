"""Steady state by FP64 defect correction (fuss_refine) on the GPU.

The refined form must stop on the reference's own test -- the residual
||T^{n+1}-T^n||_2 of an FP64 Explicit_Step below res-threshold -- and land on
the same discrete steady state as the plain FP64 march.  A march that stops
at residual r with per-iteration contraction rho is still up to
r rho / (1 - rho) (2-norm) away from its fixed point, and so is the refined
state by its own final residual; the comparison bound is twice the sum,
with rho measured from the tail of the plain residual history, plus an
L-inf sanity bound of 1e-10 relative.  The FP32- and FP64-inner forms
converge to the same FP64 fixed point and agree to a few ulps.
"""
import numpy as np
import pytest

from fuss_amd import case as cs
from fuss_amd import cases
from tests.conftest import gpu_available

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gpu():
    if not gpu_available():
        pytest.skip("no GPU")
    from fuss_amd import gpu as g

    g.load()
    return g


def relerr(a, b):
    return float(np.max(np.abs(a - b) / np.abs(b)))


def plain_steady(gpu, c):
    """Plain FP64 march to res-threshold: (T of every block, iterations,
    final residual, contraction per iteration from the last 10% of steps)."""
    s = gpu.Solver(c)
    n, hist = s.solve()
    T = s.T()
    s.close()
    r = hist[:, 2]
    assert r[-1] <= c.params.res_threshold
    k = max(10, len(r) // 10)
    rho = (r[-1] / r[-1 - k]) ** (1.0 / k)
    return T, n, r[-1], rho


def check_close(Tr, res_ref, plain):
    Tp, _, res_plain, rho = plain
    bound = 2.0 * (res_plain + res_ref) * rho / (1.0 - rho)
    d = np.sqrt(sum(float(np.sum((a - b) ** 2)) for a, b in zip(Tr, Tp)))
    assert d <= bound, (d, bound)
    for a, b in zip(Tr, Tp):
        assert relerr(a, b) < 1e-10


def refined(gpu, c, precision, ctx_precision=None, **kw):
    s = gpu.Solver(c) if ctx_precision is None else gpu.Solver(c, precision=ctx_precision)
    rp = gpu.refine_params(precision=precision, res_threshold=c.params.res_threshold, **kw)
    r, hist = s.refine(rp)
    T = s.T()
    s.close()
    return T, r, hist


def box(n=(14, 10, 8), faces=None, T0=400.0, k=20.0, qvol=2e5):
    """Single 3-D block with a mix of wall types (Cartesian, uneven spacing)."""
    rng = np.random.default_rng(11)
    axes = [np.concatenate([[0.0], np.cumsum(rng.uniform(0.5, 1.5, m))]) * 0.01 for m in n]
    nodes = cases.nodes_from_axes(*axes)
    shape = (n[2], n[1], n[0])
    T = np.full(shape, T0)
    blk = cs.Block(tuple(n), nodes, T, np.ones(shape), np.full(shape, qvol))
    faces = faces or [(302, cases.R(Tw=500.0)), (304, cases.R(Tref=300.0, eps=0.8)),
                      (303, cases.R(hconv=20.0, Tref=300.0)), (305, cases.R(hconv=10.0, Tref=280.0, eps=0.5)),
                      (301, cases.R(qw=-1e4)), (0, cases.R())]
    bcs = [cases.bclist(*cases.single_block_bc(1, tuple(n), faces))]
    p = cs.Params(n_rk=3, vnn=0.8, prim=True, irs_beta=0.0, mgl=1, level_iter=[10**9], res_threshold=1e-10)
    return cs.Case([blk], bcs, cases.const_table(k, 8000.0, 500.0), p, [], [np.full(bcs[0].n, -1, np.int32)])


@pytest.mark.parametrize("n_rk", [3, 4])
def test_cube_matches_plain(gpu, n_rk):
    c = cases.cube3d(n=16, mgl=1, n_rk=n_rk)
    plain = plain_steady(gpu, c)
    T32, r32, h32 = refined(gpu, cases.cube3d(n=16, mgl=1, n_rk=n_rk), gpu.FUSS_FP32)
    T64, r64, _ = refined(gpu, cases.cube3d(n=16, mgl=1, n_rk=n_rk), gpu.FUSS_FP64)
    assert r32.converged and r64.converged
    assert r32.res <= 1e-10 and h32[-1] == r32.res
    assert r32.outer < 20 and plain[1] > 10000
    check_close(T32, r32.res, plain)
    check_close(T64, r64.res, plain)
    assert relerr(T32[0], T64[0]) < 1e-14


def test_cube_matches_oracle(gpu, oracle_mod):
    """The oracle's plain march to 1e-10 (the reference algorithm on the CPU)."""
    c = cases.cube3d(n=10, mgl=1)
    o = oracle_mod.Oracle(c)
    n, hist = o.run()[:2]
    r_o = hist[:, 2]
    assert r_o[-1] <= 1e-10
    k = max(10, len(r_o) // 10)
    rho = (r_o[-1] / r_o[-1 - k]) ** (1.0 / k)
    T32, r, _ = refined(gpu, cases.cube3d(n=10, mgl=1), gpu.FUSS_FP32)
    assert r.converged
    check_close(T32, r.res, ([o.get("T")], n, r_o[-1], rho))


def test_plate2d_matches_plain(gpu):
    """2-D plate with 302/301 walls (the 302 derivative enters D)."""
    plain = plain_steady(gpu, cases.plate2d(48))
    T32, r, _ = refined(gpu, cases.plate2d(48), gpu.FUSS_FP32)
    assert r.converged
    check_close(T32, r.res, plain)


def test_mixed_walls_matches_plain(gpu):
    """302, 304 and 305 radiative, 303 convective, 301 flux and null walls."""
    plain = plain_steady(gpu, box())
    T32, r, _ = refined(gpu, box(), gpu.FUSS_FP32)
    assert r.converged
    check_close(T32, r.res, plain)


def test_inner_irs_and_limited_depth(gpu):
    """IRS in the smoother, and a 3-level V-cycle whose coarsest level is
    smoothed rather than solved directly: slower, same fixed point."""
    plain = plain_steady(gpu, cases.cube3d(n=16, mgl=1))
    T1, r1, _ = refined(gpu, cases.cube3d(n=16, mgl=1), gpu.FUSS_FP32, irs_beta=0.5, lam_hi=1.2, lam_lo=0.2)
    T2, r2, _ = refined(gpu, cases.cube3d(n=16, mgl=1), gpu.FUSS_FP32, max_levels=3, ncoarse=40)
    assert r1.converged and r2.converged
    assert r2.levels == 3
    check_close(T1, r1.res, plain)
    check_close(T2, r2.res, plain)


def test_independent_blocks(gpu):
    """Several unconnected cubes in one context (the parameter sweep layout)."""
    plain = plain_steady(gpu, cases.sweep(ncase=3, lo=8, hi=16, mgl=1))
    T32, r, _ = refined(gpu, cases.sweep(ncase=3, lo=8, hi=16, mgl=1), gpu.FUSS_FP32)
    assert r.converged
    check_close(T32, r.res, plain)


def test_connections_rejected(gpu):
    s = gpu.Solver(cases.lattice(nl=2, lo=4, hi=6, time_accurate=False))
    with pytest.raises(gpu.FussError) as ei:
        s.refine(gpu.refine_params())
    assert ei.value.code == 5  # FUSS_EUNSUPPORTED
    s.close()


@pytest.mark.parametrize("precision", ["FUSS_FP32", "FUSS_FP64"])
def test_three_levels_cg_coarse(gpu, precision):
    """A 3-level V-cycle whose coarsest level (8^3 of a 32^3 cube) is solved
    by conjugate gradients converges to the full hierarchy's fixed point in
    about as many outer iterations."""
    n = 32
    Tf, rf, _ = refined(gpu, cases.cube3d(n=n, mgl=1), getattr(gpu, precision))
    Tc, rc, hc = refined(gpu, cases.cube3d(n=n, mgl=1), getattr(gpu, precision), max_levels=3, coarse_cg=40)
    assert rf.converged and rc.converged and rc.levels == 3
    assert rc.outer <= rf.outer + 3 and hc[-1] == rc.res
    for a, b in zip(Tc, Tf):
        assert relerr(a, b) < 1e-11


def test_cg_coarse_rejects_negative(gpu):
    s = gpu.Solver(cases.cube3d(n=8, mgl=1))
    with pytest.raises(gpu.FussError):
        s.refine(gpu.refine_params(max_levels=2, coarse_cg=-1))
    s.close()
