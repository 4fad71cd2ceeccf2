"""Persistent launches (fuss_persistent): fuss_run's steps in one cooperative
launch with grid-wide barriers and device-side stop tests.

The persistent path uses the step loop's arithmetic phase by phase (dt,
fused fluxes + BC values + update, mirror ghosts, the fixed-order residual
reduction), so it must stop at the same iteration with the same state,
residual history, clock and status as the captured-graph path, bit for bit;
levels it does not cover (IRS, connections) keep the graph path.
"""
import numpy as np
import pytest

from fuss_amd import cases
from tests import cases3d as c3
from tests.golden import fixtures

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gpu():
    from tests.conftest import gpu_available

    if not gpu_available():
        pytest.skip("no GPU")
    from fuss_amd import gpu as g

    g.load()
    return g


def run(gpu, c, persistent, n, precision=None, thr=None, tthr=None, itermax=None):
    s = gpu.Solver(c) if precision is None else gpu.Solver(c, precision=precision)
    s.persistent(persistent)
    k, hr, ht, r = s.run_level(1, n, itermax or n, c.params.res_threshold if thr is None else thr,
                               c.params.time_threshold if tthr is None else tthr)
    T = np.concatenate([x.ravel() for x in s.T()])
    clock = s.clock()
    s.close()
    return k, hr, ht, T, clock, r


def same(a, b):
    ka, hra, hta, Ta, ca, ra = a
    kb, hrb, htb, Tb, cb, rb = b
    assert ka == kb, (ka, kb)
    assert np.array_equal(hra, hrb) and np.array_equal(hta, htb)
    assert np.array_equal(Ta, Tb)
    assert ca == cb
    assert ra.res_sq == rb.res_sq and ra.dtglobal == rb.dtglobal


@pytest.mark.parametrize("name", ["oscillating_Rod", "convective_Plate", "radiative_Rod", "qvol_Plate"])
def test_persistent_equals_graph_reference_cases(gpu, name):
    """The reference's cases: the time-accurate periodic-series rod to its
    time limit; the steady plates and the radiative rod (primitive and
    conservative variables, 301/302/303/304 walls, qvol) with IRS off
    (vnn 0.5), 3000 steps."""
    c, _ = fixtures.load(name)
    if name != "oscillating_Rod":
        c.params.irs_beta = 0.0
        c.params.vnn = 0.5
    n = 10**6 if name == "oscillating_Rod" else 3000
    mk = lambda: c  # noqa: E731
    a = run(gpu, mk(), True, n)
    b = run(gpu, mk(), False, n)
    print(f"{name}: {a[0]} steps", flush=True)
    same(a, b)


def test_persistent_plate_to_convergence(gpu):
    """The 2-D parity plate (64^2) to its residual 1e-10: same stop iteration,
    history and field as the graph path."""
    c = cases.plate2d(64)
    a = run(gpu, c, True, 10**6)
    b = run(gpu, cases.plate2d(64), False, 10**6)
    assert a[1][-1] <= 1e-10
    same(a, b)


@pytest.mark.parametrize("name", ["box_walls", "affine_walls", "box_rk4_transient", "sweep_noirs"])
def test_persistent_3d(gpu, oracle_mod, name):
    """3-D levels: a stretched box with 302/303/304/305/301/300 walls, the
    sheared block (tangential weights), a time-accurate four-stage box, and
    independent blocks of different sizes (one launch over all their tiles),
    200 steps.  The persistent path runs the generic tile stage kernel where
    the graph path runs the column kernels of 3-D levels (same arithmetic up
    to the regrouping of the per-axis metric products), so both are checked
    against each other and against the oracle to 1e-12."""
    from tests.test_gpu_parity import relerr

    if name == "box_walls":
        mk = lambda: c3.box(faces=[(302, cases.R(Tw=600.0)), (305, cases.R(eps=0.7, hconv=35.0, Tref=300.0)),  # noqa: E731
                                   (303, cases.R(hconv=20.0, Tref=300.0)), (304, cases.R(eps=0.8, Tref=300.0)),
                                   (301, cases.R(qw=-1e4)), (300, cases.R())], T0=520.0)
    elif name == "affine_walls":
        mk = lambda: c3.affine_quadratic(n_rk=3, faces=[(302, cases.R(Tw=480.0)), (303, cases.R(hconv=40.0, Tref=300.0)),  # noqa: E731
                                                        (304, cases.R(eps=0.8, Tref=300.0)), (301, cases.R()),
                                                        (301, cases.R(qw=2e3)), (300, cases.R())])
    elif name == "box_rk4_transient":
        mk = lambda: c3.box(n=(9, 8, 7), n_rk=4, time_accurate=True)  # noqa: E731
    else:
        mk = lambda: cases.sweep(ncase=3, lo=12, hi=28, n_rk=3, irs_beta=0.0, vnn=0.5, mgl=1, seed=7)  # noqa: E731
    a = run(gpu, mk(), True, 200, thr=-1.0, tthr=1e300)
    b = run(gpu, mk(), False, 200, thr=-1.0, tthr=1e300)
    assert a[0] == b[0] == 200
    assert relerr(a[3], b[3]) < 1e-12
    np.testing.assert_allclose(a[1], b[1], rtol=1e-9)
    np.testing.assert_allclose(a[2], b[2], rtol=1e-13)
    c = mk()
    o = oracle_mod.Oracle(c)
    for _ in range(200):
        o.step(1)
    To = np.concatenate([x.ravel() for x in o.T()])
    assert relerr(a[3], To) < 1e-12


def test_persistent_chunks_and_limits(gpu):
    """Step limits below and across launch chunks, and a second call
    continuing the first: the same as the graph path."""
    for n in (1, 5, 37):
        same(run(gpu, cases.plate2d(32), True, n, thr=-1.0), run(gpu, cases.plate2d(32), False, n, thr=-1.0))
    s = gpu.Solver(cases.plate2d(32))
    s.persistent(True)
    s.run_level(1, 40, 40, -1.0, 1e300)
    k, hr, _, _ = s.run_level(1, 40, 40, -1.0, 1e300)
    T1 = s.get("T")
    s.close()
    s = gpu.Solver(cases.plate2d(32))
    k2, hr2, _, _ = s.run_level(1, 80, 80, -1.0, 1e300)
    assert np.array_equal(s.get("T"), T1) and np.array_equal(hr2[40:], hr)
    s.close()


def test_persistent_fp32(gpu):
    """FP32 contexts: the persistent FP32 steps equal the FP32 graph path."""
    a = run(gpu, cases.plate2d(48), True, 300, precision=gpu.FUSS_FP32, thr=-1.0)
    b = run(gpu, cases.plate2d(48), False, 300, precision=gpu.FUSS_FP32, thr=-1.0)
    same(a, b)


def test_persistent_state_check(gpu, oracle_mod):
    """A failing state check ends the launch after the failing step with
    FUSS_ESTATE and the reference's first failing cell."""
    c = c3.drain_plate(cells=(5, 3), n_rk=3)
    s = gpu.Solver(c)
    s.persistent(True)
    with pytest.raises(gpu.FussError) as ei:
        s.run_level(1, 50, 50, -1.0, 1e300, hist=False)
    assert ei.value.code == gpu.FUSS_ESTATE and ei.value.bijk == (1, 1, 3, 1)
    s.close()


def test_persistent_falls_back(gpu):
    """IRS levels and connected lattices keep the graph path (same results
    with the switch on or off)."""
    for mk in (lambda: cases.plate2d(32, irs_beta=0.5), lambda: cases.lattice(nl=2, lo=4, hi=8)):
        same(run(gpu, mk(), True, 60, thr=-1.0, tthr=1e300), run(gpu, mk(), False, 60, thr=-1.0, tthr=1e300))


def sheared_plate(n=24, m=20):
    """2-D plate on a sheared grid (tangential weights), walls on four sides."""
    from fuss_amd import case as cs

    x = np.arange(n + 1) * 0.01
    y = np.arange(m + 1) * 0.012
    nodes = cases.nodes_from_axes(x, y, np.array([0.0, 0.01]))
    nodes[..., 0] += 0.3 * nodes[..., 1]
    rng = np.random.default_rng(4)
    T = 400.0 + 30.0 * rng.uniform(-1, 1, size=(1, m, n))
    faces = [(302, cases.R(Tw=520.0)), (303, cases.R(hconv=40.0, Tref=300.0)), (300, cases.R()),
             (304, cases.R(eps=0.8, Tref=300.0)), (0, cases.R()), (0, cases.R())]
    bcs = [cases.bclist(*cases.single_block_bc(1, (n, m, 1), faces))]
    p = cs.Params(n_rk=3, vnn=0.5, prim=True, irs_beta=0.0, mgl=1, level_iter=[10**9], res_threshold=1e-10)
    return cs.Case([cs.Block((n, m, 1), nodes, T, np.ones_like(T), np.zeros_like(T))], bcs,
                   cases.const_table(20.0, 8000.0, 500.0), p, [], [np.full(bcs[0].n, -1, np.int32)])


FUSED_CASES = {
    "plate2d_64": lambda: cases.plate2d(64),
    "plate2d_rk4": lambda: cases.plate2d(40, n_rk=4),
    "oscillating_Rod": lambda: fixtures.load("oscillating_Rod")[0],
    "convective_noirs": lambda: _noirs("convective_Plate"),
    "radiative_noirs": lambda: _noirs("radiative_Rod"),
    "qvol_cons_noirs": lambda: _noirs("qvol_Plate"),
    "sheared_plate": sheared_plate,
}


def _noirs(name):
    c, _ = fixtures.load(name)
    c.params.irs_beta = 0.0
    c.params.vnn = 0.5
    return c


@pytest.mark.parametrize("graphs", [False, True])
@pytest.mark.parametrize("name", list(FUSED_CASES))
def test_fused_stages_equal_split_passes(gpu, name, graphs):
    """2-D levels without IRS or connections: one fused launch per stage
    (BC values in place, mirror ghosts with the new state, dt in the first
    stage of steady steps) gives the split passes' steps bit for bit --
    state, residual history, clock, dtglobal -- in the step loop and on
    captured graphs."""
    out = []
    for fused in (True, False):
        c = FUSED_CASES[name]()
        s = gpu.Solver(c)
        s.fused(fused)
        s.graphs(graphs)
        k, hr, ht, r = s.run_level(1, 150, 150, -1.0, c.params.time_threshold)
        out.append((k, hr, ht, np.concatenate([x.ravel() for x in s.T()]), s.clock(), r))
        s.close()
    same(out[0], out[1])


@pytest.mark.parametrize("mode", ["fused", "split", "persistent"])
def test_c1_plate128_to_convergence(gpu, mode):
    """The 2-D parity configuration (128 x 128 plate, RK3, local dt, vnn 0.8,
    no IRS) from its seeded state to residual 1e-10 on the GPU: the oracle's
    stop iteration (fixture tests/golden/plate2d_128_oracle.npz, made by
    tests/golden/make_plate128.py), its field to 1e-12 and its residual
    history (every 500th iteration) to 1e-9."""
    import os

    from tests.test_gpu_parity import relerr

    z = np.load(os.path.join(os.path.dirname(__file__), "golden", "plate2d_128_oracle.npz"), allow_pickle=False)
    c = cases.plate2d(128)
    s = gpu.Solver(c)
    s.fused(mode != "split")
    s.persistent(mode == "persistent")
    n, hist = s.solve()
    T = s.get("T")
    s.close()
    assert n == int(z["stop_iter"]), (n, int(z["stop_iter"]))
    assert relerr(T, z["T"]) < 1e-12
    np.testing.assert_allclose(hist[z["res_iters"] - 1, 2], z["res"], rtol=1e-9)
    assert abs(hist[-1, 2] - float(z["res_last"])) <= 1e-9 * float(z["res_last"])
