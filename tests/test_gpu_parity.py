"""Parity of the HIP path (libfuss_gpu.so, FP64) with the oracle and with the
reference's stored outputs.

Tolerances: the bar for the plain FP64 path is 1e-12 relative (L-inf on
T).  Metrics are computed on the device with the reference's
operation order and FP contraction off, so they must agree bit for bit.
The interior fluxes regroup A * kappa * (grad . n) as kappa * sum_c G_c w_c
with precomputed face weights w (a different rounding, ~1 ulp per face), so
fields agree to ~1e-15 per step rather than bitwise.
"""
import numpy as np
import pytest

from fuss_amd import cases
from fuss_amd import formats as fm
from tests.conftest import gpu_available
from tests import known_answers
from tests.golden import fixtures
from tests.test_oracle_golden import check_wall

pytestmark = pytest.mark.gpu
RTOL = 1e-12
ALL = list(fixtures.NAMES)
STEADY = ["temperature_Plate", "qvol_Plate", "convective_Plate", "radiative_Rod", "chimera"]


@pytest.fixture(scope="module")
def gpu():
    if not gpu_available():
        pytest.skip("no GPU")
    from fuss_amd import gpu as g

    g.load()
    return g


def relerr(a, b):
    return float(np.max(np.abs(a - b) / np.abs(b)))


@pytest.mark.parametrize("name", ["temperature_Plate", "multimat_Plate", "mg_irs"])
def test_metrics_bitwise(gpu, oracle_mod, name):
    c, _ = fixtures.load(name)
    s = gpu.Solver(c)
    o = oracle_mod.Oracle(c)
    for lv in range(1, c.params.mgl + 1):
        for b in range(1, c.nb + 1):
            M, vol, faces = s.metrics(lv, b)
            Mo, dlo, volo, faceso = o.metrics(lv, b)
            ni, nj, nk = s.dims(lv, b)
            sl = (slice(2, 2 + nk), slice(2, 2 + nj), slice(2, 2 + ni))
            np.testing.assert_array_equal(M, Mo[sl])
            np.testing.assert_array_equal(vol, volo[sl])
            for d in range(3):
                np.testing.assert_array_equal(faces[d][0], faceso[d][0])
                np.testing.assert_array_equal(faces[d][1], faceso[d][1])
    s.close()


def steps_vs_oracle(gpu, oracle_mod, c, nstep=20):
    """nstep Explicit_Steps on the coarsest level (and 5 on the next finer one,
    via prolongation), field and residual against the oracle."""
    s = gpu.Solver(c)
    o = oracle_mod.Oracle(c)
    lv = c.params.mgl
    for n in range(nstep):
        r = s.explicit_step(lv)
        ro = o.step(lv)
        if n == 0:
            for b in range(c.nb):
                assert relerr(s.get("T", lv, b + 1), o.get("T", lv, b + 1)) < 1e-14
        assert abs(np.sqrt(r.res_sq) - np.sqrt(ro)) <= 1e-9 * np.sqrt(ro) + 1e-12
    for b in range(c.nb):
        assert relerr(s.get("T", lv, b + 1), o.get("T", lv, b + 1)) < RTOL
    t, it = s.clock(lv)
    assert it == nstep and abs(t - o.time(lv)) <= 1e-14 * max(1.0, abs(t))
    if lv > 1:
        s.prolong(lv - 1)
        o.prolong(lv - 1)
        for n in range(5):
            s.explicit_step(lv - 1)
            o.step(lv - 1)
        for b in range(c.nb):
            assert relerr(s.get("T", lv - 1, b + 1), o.get("T", lv - 1, b + 1)) < RTOL
    s.close()


@pytest.mark.parametrize("name", ALL)
def test_steps_vs_oracle(gpu, oracle_mod, name):
    c, _ = fixtures.load(name)
    steps_vs_oracle(gpu, oracle_mod, c)


def tridiag(c, beta=None):
    """The case with the line-tridiagonal IRS (and IRS on, beta given)."""
    if beta is not None:
        c.params.irs_beta = beta
    c.params.irs_mode = 1
    return c


def with_type(c, old, new):
    """The case with every BC entry of type `old` changed to `new`."""
    for bcl in c.bcs:
        bcl.ib[bcl.ib[:, 5] == old, 5] = new
    return c


SYNTH = {
    # 3-D orthogonal stretched cube: 303 walls, qvol, k(T) table, IRS, 3 levels
    "cube3d_mg3_irs": lambda: cases.cube3d(24, mgl=3),
    # 3-D curvilinear 8-block lattice: connections, 3 materials, 301/302/304,
    # time-accurate global dt, tangential flux weights
    "lattice_transient": lambda: cases.lattice(nl=2, lo=4, hi=8),
    # the same lattice, conservative variables, IRS, 2 multigrid levels
    "lattice_cons_irs_mg2": lambda: cases.lattice(nl=2, lo=4, hi=8, prim=False, irs_beta=0.5,
                                                  time_accurate=False, mgl=2),
    # 3-D lattice with uneven tiles (dims not multiples of the 64x4 tile) and 3 levels
    "lattice_mg3": lambda: cases.lattice(nl=2, lo=8, hi=20, irs_beta=0.5, time_accurate=False, mgl=3),
    # 2-D plate parity configuration (reduced size)
    "plate2d_64": lambda: cases.plate2d(64),
    # four-stage (Jameson) extension: the cube with IRS, and the transient
    # lattice in conservative variables
    "cube3d_rk4_irs": lambda: cases.cube3d(24, mgl=1, n_rk=4),
    # three independent rectilinear blocks of different sizes and materials
    # (per-axis metric factors per block), four-stage RK with IRS
    "sweep3_rk4_irs": lambda: cases.sweep(ncase=3, lo=12, hi=28, n_rk=4, mgl=1, seed=7),
    "lattice_rk4_cons": lambda: cases.lattice(nl=2, lo=4, hi=8, n_rk=4, prim=False),
    # a row of three curvilinear blocks joined on i faces only: the fused
    # x-y-z IRS pass with donor x ghosts (steady, local dt)
    "xrow_irs": lambda: cases.lattice(nl=(3, 1, 1), lo=6, hi=20, irs_beta=0.5, time_accurate=False),
    # line-tridiagonal IRS extension (FUSS_IRS_TRIDIAG): single block, the
    # connected lattice (donor ghosts between directions, 2 levels), uneven
    # line lengths (not multiples of the 16-cell chunks), and 2-D
    "cube3d_tridiag": lambda: tridiag(cases.cube3d(24, mgl=1, n_rk=4)),
    "lattice_tridiag_mg2": lambda: tridiag(cases.lattice(nl=2, lo=4, hi=8, prim=False, irs_beta=0.5,
                                                         time_accurate=False, mgl=2)),
    "lattice_tridiag_mg3": lambda: tridiag(cases.lattice(nl=2, lo=8, hi=20, irs_beta=0.5, time_accurate=False,
                                                         mgl=3)),
    "plate2d_tridiag": lambda: tridiag(cases.plate2d(40), beta=0.6),
    # the 2-D plate with its 301 walls turned into 103 coupling entries
    # (ext_flux 0: no flux, mirrored ghost)
    "plate2d_103": lambda: with_type(cases.plate2d(32), 301, 103),
}


@pytest.mark.parametrize("name", list(SYNTH))
def test_synthetic_vs_oracle(gpu, oracle_mod, name):
    steps_vs_oracle(gpu, oracle_mod, SYNTH[name]())


@pytest.mark.parametrize("name", STEADY)
def test_steady_to_convergence(gpu, name):
    """Run to res-threshold on the GPU: same stop iteration as the reference
    and its stored field to 1e-12."""
    c, z = fixtures.load(name)
    s = gpu.Solver(c)
    n, hist = s.solve()
    assert n == int(z["stop_iter"]), (n, int(z["stop_iter"]))
    for b in range(c.nb):
        err = relerr(s.get("T", 1, b + 1), z[f"ref_T{b}"])
        print(f"{name}: stop {n}, block {b+1} max rel err vs reference {err:.3e}", flush=True)
        assert err < RTOL
    Tw, qw = s.wall(1)
    check_wall(c, z, Tw, qw)
    # the published known answer the reference's verify.py checks (< 1 %)
    known_answers.check_steady(c, name, [s.get("T", 1, b + 1) for b in range(c.nb)], Tw)
    s.close()


@pytest.mark.parametrize("name", ["box_walls", "lattice_transient", "plate2d_64", "convective"])
def test_wall_vs_oracle(gpu, oracle_mod, name):
    """fuss_wall (Extract_Properties_From_BC on the device) against the
    oracle after 8 steps: walls 301 (with a time series on the lattice),
    302, 303, 304, 305 in 3-D, curvilinear multi-block and 2-D."""
    from tests.test_refine import box

    if name == "box_walls":
        c = box()
    elif name == "convective":
        c, _ = fixtures.load("convective_Plate")
    else:
        c = SYNTH[name]()
    s = gpu.Solver(c)
    o = oracle_mod.Oracle(c)
    for _ in range(8):
        s.explicit_step(1)
        o.step(1)
    Tw, qw = s.wall(1)
    # each call allocates its output on the device: repeated calls return
    # the same values (the output's zero fill must finish before the wall
    # kernel writes it)
    for _ in range(10):
        Tw2, qw2 = s.wall(1)
        assert np.array_equal(Tw2, Tw) and np.array_equal(qw2, qw)
    Two, qwo = o.wall(1)
    walls = np.isin(c.bcs[0].ib[:, 5], fm.WALL_TYPES)
    assert walls.any() and not Tw[~walls].any() and not qw[~walls].any()
    assert np.max(np.abs(Tw[walls] - Two[walls]) / np.abs(Two[walls])) < RTOL
    assert np.max(np.abs(qw[walls] - qwo[walls])) <= RTOL * np.max(np.abs(qwo[walls]))
    s.close()


def test_multimat_transient(gpu):
    c, z = fixtures.load("multimat_Plate")
    s = gpu.Solver(c)
    it = 0
    for q, target in enumerate(z["snap_iter"]):
        while it < target:
            s.explicit_step(1, compute_residual=False)
            it += 1
        t, _ = s.clock(1)
        assert abs(t - float(z[f"snap{q}_time"])) <= 1e-13 * abs(t)
        for b in range(c.nb):
            assert relerr(s.get("T", 1, b + 1), z[f"snap{q}_T{b}"]) < RTOL
        assert known_answers.multimat_error(c, s.get("T", 1, 2), t) < 0.02
    s.close()


def test_oscillating_probes(gpu):
    c, z = fixtures.load("oscillating_Rod")
    s = gpu.Solver(c)
    p1, p2 = z["p1"], z["p2"]
    got = []
    for row in range(p1.shape[0]):
        r = s.explicit_step(1, compute_residual=False)
        v = s.probe(z["probes"])
        assert abs(r.time - p1[row, 0]) <= 5e-14 * p1[row, 0]
        assert abs(v[0] - p1[row, 1]) <= RTOL * p1[row, 1]
        assert abs(v[1] - p2[row, 1]) <= RTOL * p2[row, 1]
        got.append(v)
    got = np.array(got)
    assert known_answers.oscillating_error(got[:, 0], got[:, 1]) < 0.01
    s.close()


def test_mg_irs_trajectory(gpu):
    """All 80603 residuals of the 2-level multigrid + IRS + connection run."""
    from tests.test_oracle_golden import residual_tolerance

    c, z = fixtures.load("mg_irs")
    ref = z["res_hist"]
    s = gpu.Solver(c)
    n, hist = s.solve(max_iters=ref.size + 10)
    assert n == ref.size
    tol = residual_tolerance(ref, 2000.0, c.ncell)
    bad = np.abs(hist[:, 2] - ref) > tol
    print(f"mg_irs: {n} iterations, max |dres|/tol = {np.max(np.abs(hist[:,2]-ref)/tol):.3f}", flush=True)
    assert not bad.any(), f"first mismatch at iteration {np.argmax(bad)+1}"
    s.close()


@pytest.mark.parametrize("name", ["convective_Plate", "oscillating_Rod"])
def test_graph_run_equals_step_loop(gpu, name):
    """fuss_run replayed from captured graphs (device-side stop tests and
    time series) stops at the same step with the same state, bit for bit, as
    the step-by-step loop."""
    c, _ = fixtures.load(name)
    out = []
    for graphs in (True, False):
        s = gpu.Solver(c)
        s.graphs(graphs)
        k, hr, ht, r = s.run_level(1, 500, 500 if name == "oscillating_Rod" else 10**9, c.params.res_threshold,
                                   c.params.time_threshold)
        out.append((k, hr.copy(), ht.copy(), s.get("T"), s.clock()))
        s.close()
    (k1, hr1, ht1, T1, c1), (k2, hr2, ht2, T2, c2) = out
    assert k1 == k2 and c1 == c2
    assert np.array_equal(hr1, hr2) and np.array_equal(ht1, ht2)
    assert np.array_equal(T1, T2)


@pytest.mark.parametrize("rect", [False, True])
def test_graph_run_fused_dt(gpu, rect):
    """The cube on captured graphs equals the step loop bit for bit: with the
    stored metrics the fused x-y-z IRS pass leaves the next step's dt behind;
    with the per-axis factors (steady, primitive) the first stage computes it."""
    out = []
    for graphs in (True, False):
        s = gpu.Solver(SYNTH["cube3d_rk4_irs"]())
        s.rect(rect)
        s.graphs(graphs)
        k, hr, _, _ = s.run_level(1, 60, 60, -1.0, 1e300)
        out.append((k, hr.copy(), s.get("T")))
        s.close()
    assert out[0][0] == out[1][0] == 60
    assert np.array_equal(out[0][1], out[1][1]) and np.array_equal(out[0][2], out[1][2])


@pytest.mark.parametrize("steps", [20, 7, 44, 6])
def test_graph_chunk_sizes(gpu, steps):
    """Requests below one 30-step chunk capture a chunk of the request
    rounded up to a multiple of 3 (the extra steps are device no-ops), and a
    request that is not a multiple of 30 ends inside a chunk: the state,
    clock and residual history equal the step loop's bit for bit, and a
    second call with the same size replays the kept graph."""
    out = []
    for graphs in (True, False):
        s = gpu.Solver(SYNTH["lattice_rk4_cons"]())
        s.graphs(graphs)
        rows = []
        for _ in range(2):
            k, hr, ht, _ = s.run_level(1, steps, steps, -1.0, 1e300)
            rows.append((k, hr.copy(), ht.copy()))
        out.append((rows, np.concatenate([T.ravel() for T in s.T()]), s.clock()))
        s.close()
    (r1, T1, c1), (r2, T2, c2) = out
    assert [x[0] for x in r1] == [x[0] for x in r2] == [steps, steps]
    for a, b in zip(r1, r2):
        assert np.array_equal(a[1], b[1]) and np.array_equal(a[2], b[2])
    assert np.array_equal(T1, T2) and c1 == c2


@pytest.mark.parametrize("rect", [False, True])
def test_dtglobal_vs_oracle(gpu, oracle_mod, rect):
    """dtglobal of each step (k_dt, the fused IRS pass's next-step dt, or the
    first stage's dt) against the oracle's Compute_dt: bit for bit on the
    first step (same state), to rounding after."""
    c = SYNTH["cube3d_rk4_irs"]()
    s = gpu.Solver(c)
    s.rect(rect)
    o = oracle_mod.Oracle(c)
    for n in range(6):
        r = s.explicit_step(1)
        o.step(1)
        if n == 0:
            assert r.dtglobal == o.dtglobal(1)
        else:
            assert abs(r.dtglobal - o.dtglobal(1)) <= 1e-13 * o.dtglobal(1)
    s.close()


def test_set_T_between_steps(gpu, oracle_mod):
    """A state set between steps (restart) is stepped with its own dt: the
    next step's dt left by the fused IRS pass is dropped."""
    c = SYNTH["cube3d_rk4_irs"]()
    s = gpu.Solver(c)
    o = oracle_mod.Oracle(c)
    for _ in range(3):
        s.explicit_step(1)
        o.step(1)
    T = o.get("T") + 5.0 * np.sin(np.arange(o.get("T").size)).reshape(o.get("T").shape)
    s.set_T(T)
    o.set("T", T)
    for _ in range(3):
        s.explicit_step(1)
        o.step(1)
    assert relerr(s.get("T"), o.get("T")) < RTOL
    s.close()


def test_rect_factors_active(gpu):
    """The setup check accepts the per-axis factors on a stretched orthogonal
    cube (every level) and rejects them on the perturbed lattice."""
    c = SYNTH["cube3d_mg3_irs"]()
    s = gpu.Solver(c)
    for lv in range(1, c.params.mgl + 1):
        w, d, dev = s.rect_info(lv)
        assert w and d and 0.0 <= dev <= 1e-14, (lv, w, d, dev)
    s.rect(False)
    assert s.rect_info(1)[:2] == (False, False)
    s.close()
    c = SYNTH["lattice_transient"]()
    s = gpu.Solver(c)
    w, d, dev = s.rect_info(1)
    assert not w and not d and dev == -1.0
    s.close()


@pytest.mark.parametrize("graphs", [False, True])
def test_rect_vs_stored_metrics(gpu, oracle_mod, graphs):
    """Stage passes with the per-axis factors and with the stored 3-D metric
    arrays agree to rounding (the factors differ from the stored weights by a
    few ulp), and both meet the oracle's field to 1e-12 after 40 steps."""
    c = SYNTH["cube3d_rk4_irs"]()
    o = oracle_mod.Oracle(c)
    for _ in range(40):
        o.step(1)
    out = []
    for rect in (True, False):
        s = gpu.Solver(c)
        s.rect(rect)
        s.graphs(graphs)
        if graphs:
            s.run_level(1, 40, 40, -1.0, 1e300)
        else:
            for _ in range(40):
                s.explicit_step(1)
        out.append(s.get("T"))
        s.close()
    assert relerr(out[0], out[1]) < 1e-13
    for T in out:
        assert relerr(T, o.get("T")) < RTOL


@pytest.mark.parametrize("where", ["wall", "interior", "below", "all_huge"])
@pytest.mark.parametrize("graphs", [False, True])
def test_table_range_error(gpu, oracle_mod, where, graphs):
    """A temperature outside the property table rows (T in [1, 2000) K here)
    stops the step with the table status, as the table lookups of the
    reference stop (Lib_Solid.f90:10-98), in the oracle and on the device:
    on a wall cell (the wall-flux pass flags it), an interior cell, below the
    first row, and a field far beyond the int range of the row index."""
    c = SYNTH["cube3d_rk4_irs"]()
    s = gpu.Solver(c)
    s.graphs(graphs)
    o = oracle_mod.Oracle(c)
    T = o.get("T").copy()
    if where == "wall":
        T[5, 7, 0] = 2500.0
    elif where == "interior":
        T[12, 12, 12] = 2500.0
    elif where == "below":
        T[12, 12, 12] = 0.5
    else:
        T[:] = 1e12
    s.set_T(T)
    o.set("T", T)
    with pytest.raises(oracle_mod.OracleError, match="status 3"):
        o.step(1)
    with pytest.raises(gpu.FussError) as ei:
        if graphs:
            s.run_level(1, 40, 40, -1.0, 1e300)
        else:
            s.explicit_step(1)
    assert ei.value.code == gpu.FUSS_ETABLE
    s.close()
    # the device runs the next context normally
    s = gpu.Solver(c)
    o = oracle_mod.Oracle(c)
    s.explicit_step(1)
    o.step(1)
    assert relerr(s.get("T"), o.get("T")) < RTOL
    s.close()


def test_graph_run_tridiag(gpu):
    """fuss_run with the tridiagonal IRS on captured graphs (its tables are
    uploaded before the capture) equals the step loop bit for bit."""
    out = []
    for graphs in (True, False):
        s = gpu.Solver(SYNTH["lattice_tridiag_mg3"]())
        s.graphs(graphs)
        k, hr, _, _ = s.run_level(3, 60, 60, -1.0, 1e300)
        out.append((k, hr.copy(), np.concatenate([T.ravel() for T in s.T(3)])))
        s.close()
    assert out[0][0] == out[1][0] == 60
    assert np.array_equal(out[0][1], out[1][1]) and np.array_equal(out[0][2], out[1][2])


@pytest.mark.parametrize("name", ["cube3d_rk4_irs", "lattice_transient"])
def test_autotune_keeps_results(gpu, name):
    """fuss_autotune changes tile shapes and kernels only: the state after
    the same steps is bit-identical, the residual equal to rounding."""
    out = []
    for tune in (False, True):
        s = gpu.Solver(SYNTH[name]())
        if tune:
            kc, row, ms = s.autotune(1, steps=2)
            assert kc in (0, 4, 8, 16, 32, 64) and ms >= 0.0
            tl = s.tiles(1)
            assert tl[0] == kc and tl[2] in ((0,) if kc == 0 else (1, 4, 8))
            assert tl[3] in ((0,) if kc == 0 else (4, 8))
        r = [s.explicit_step(1).res_sq for _ in range(6)]
        out.append((np.concatenate([T.ravel() for T in s.T()]), np.array(r)))
        s.close()
    assert np.array_equal(out[0][0], out[1][0])
    np.testing.assert_allclose(out[0][1], out[1][1], rtol=1e-13)


@pytest.mark.parametrize("rect", [False, True])
@pytest.mark.parametrize("name", ["cube3d_rk4_irs", "sweep3_rk4_irs", "cube3d_mg3_irs", "cube3d_cons", "xrow_irs"])
def test_column_rows_keep_results(gpu, name, rect):
    """The LDS-ring stage kernel and the fused x-y-z IRS pass on 8-row
    columns (fuss_set_tiles) give the 4-row columns' state bit for bit at
    each column height (ragged blocks: rows past nj masked), dtglobal bit for
    bit (the first stage's dt or the IRS pass's next-step dt, one minimum per
    8-row column), and the residual to rounding (its partial sums group by
    column).  Without IRS (conservative, 20^3: ragged columns) the stage
    kernel applies the update and writes the residual partials."""
    c = cases.cube3d(20, mgl=1, irs_beta=0.0, prim=False, vnn=0.8) if name == "cube3d_cons" else SYNTH[name]()
    out = []
    for kc, st, rows in ((8, 4, 4), (4, 8, 8), (16, 8, 4), (32, 4, 8), (64, 8, 8)):
        s = gpu.Solver(c)
        s.rect(rect)
        s.set_tiles(1, kc=kc, row_march=0, stage_rows=st, irs_rows=rows)
        assert s.tiles(1) == (kc, 0, st, rows)
        rs = [s.explicit_step(1) for _ in range(5)]
        out.append((np.concatenate([T.ravel() for T in s.T()]), np.array([r.res_sq for r in rs]),
                    np.array([r.dtglobal for r in rs])))
        if rows == 8:
            with pytest.raises(gpu.FussError):
                s.set_tiles(1, kc=12)
            with pytest.raises(gpu.FussError):
                s.set_tiles(1, kc=8, irs_rows=6)
            with pytest.raises(gpu.FussError):
                s.set_tiles(1, kc=8, stage_rows=2)
        s.close()
    for T, res, dt in out[1:]:
        assert np.array_equal(T, out[0][0])
        np.testing.assert_allclose(res, out[0][1], rtol=1e-13)
        assert np.array_equal(dt, out[0][2])


def test_chimera_setup_errors(gpu):
    """A 102 entry without its donor list is an inconsistent BC list
    (FUSS_EBC at fuss_setup), and a donor list naming a cell outside its
    block or a non-102 entry is rejected when it is set (FUSS_EARG)."""
    c, _ = fixtures.load("chimera")
    bcl = c.bcs[0]
    first = min(bcl.chim)
    for bad in ("missing", "donor", "entry"):
        chim = dict(bcl.chim)
        if bad == "missing":
            del chim[first]
        elif bad == "donor":
            n1, don, vf = chim[first]
            don = don.copy()
            don[0, 1] = 10**6
            chim[first] = (n1, don, vf)
        else:
            chim[0] = chim.pop(first)  # entry 0 is a 302 wall
        c2 = fixtures.load("chimera")[0]
        c2.bcs[0].chim = chim
        with pytest.raises(gpu.FussError) as ei:
            gpu.Solver(c2)
        msg = str(ei.value)
        assert ("fuss_setup" in msg) if bad in ("missing", "donor") else ("fuss_set_chimera" in msg), msg


def test_bc_list_errors(gpu):
    """BC lists inconsistent with the blocks are rejected at fuss_setup with
    FUSS_EBC (records built on the device for levels without chimera
    entries): entries out of (block, face) order, an entry off its face, a
    face cell named twice or not at all, a block or a donor block out of
    range, a donor cell outside its block; the unmodified list sets up."""
    def fresh():
        return cases.lattice(nl=2, lo=4, hi=8, mgl=1)

    gpu.Solver(fresh()).close()
    c0 = fresh()
    ib0 = c0.bcs[0].ib
    conn = np.nonzero(ib0[:, 5] == 101)[0][0]
    wall = np.nonzero((ib0[:, 5] != 101) & (ib0[:, 4] == 1))[0][0]
    last = ib0.shape[0] - 1
    edits = {
        "order": lambda ib, rb: (ib[[last, *range(1, last), 0]], rb[[last, *range(1, last), 0]]),
        "off_face": lambda ib, rb: (_set(ib, wall, 1, 2), rb),
        "twice": lambda ib, rb: (_set(ib, wall + 1, slice(1, 4), ib[wall, 1:4]), rb),
        "missing": lambda ib, rb: (np.delete(ib, wall, axis=0), np.delete(rb, wall, axis=0)),
        "block": lambda ib, rb: (_set(ib, wall, 0, 99), rb),
        "donor_block": lambda ib, rb: (_set(ib, conn, 6, 99), rb),
        "donor_cell": lambda ib, rb: (_set(ib, conn, 7, 10**6), rb),
    }
    for name, edit in edits.items():
        c = fresh()
        ib, rb = edit(c.bcs[0].ib.copy(), c.bcs[0].rb.copy())
        c.bcs[0] = cases.bclist(ib, rb)
        c.series_id = [np.full(ib.shape[0], -1, np.int32)]
        with pytest.raises(gpu.FussError) as ei:
            gpu.Solver(c)
        assert ei.value.code == gpu.FUSS_EBC, (name, ei.value.code)


def _set(a, row, col, v):
    a[row, col] = v
    return a


def test_xrow_takes_fused_irs(gpu):
    """Blocks joined on i faces only run the fused x-y-z IRS pass (one
    IRS launch per stage, the update in it), so xrow_irs exercises its
    donor x ghosts."""
    import ctypes

    c = SYNTH["xrow_irs"]()
    s = gpu.Solver(c)
    s.L.fuss_profile(s.h, 1)
    s.explicit_step(1)
    ms, cnt = (ctypes.c_double * 8)(), (ctypes.c_int * 8)()
    s.L.fuss_profile_read(s.h, 8, ms, cnt)
    s.L.fuss_profile(s.h, 0)
    assert cnt[3] == 0 and cnt[4] == c.params.n_rk  # FUSS_K_IRS, FUSS_K_IRS_UPDATE
    s.close()


@pytest.mark.parametrize("name", ["cube3d_mg3_irs", "lattice_transient", "plate2d_64", "lattice_rk4_cons"])
def test_stage_hook_noop(gpu, name):
    """fuss_step_hook with a hook that leaves R alone: the unfused stage
    (raw R, host round trip, newstate pass) reproduces the step loop, and the
    hook sees every stage of every block with the stage's input T."""
    c = SYNTH[name]()
    lv = c.params.mgl
    out = []
    for hooked in (False, True):
        s = gpu.Solver(c)
        s.rect(False)
        calls = []

        # (an exception inside a ctypes callback is printed, not raised:
        # record, then check outside)
        def hook(level, irk, block, T, R):
            calls.append((irk, block, level, T.shape, R.shape))

        for _ in range(4):
            s.explicit_step(lv, hook=hook if hooked else None)
        out.append([s.get("T", lv, b + 1) for b in range(c.nb)])
        if hooked:
            assert len(calls) == 4 * c.params.n_rk * c.nb
            for irk, block, level, ts, rs in calls:
                assert level == lv and ts == rs == tuple(reversed(s.dims(level, block)))
            assert [q[0] for q in calls[: c.params.n_rk * c.nb]] == [irk for irk in range(1, c.params.n_rk + 1)
                                                                        for _ in range(c.nb)]
        s.close()
    for a, b in zip(out[0], out[1]):
        assert relerr(b, a) < 1e-14


def test_stage_hook_source_vs_oracle(gpu, oracle_mod):
    """A hook adding a volumetric source q0 (R -= q0 V per cell) is the same
    step as qvol + q0 in the oracle (Newstate: res = (-R/V + qvol) dt / ...)."""
    c = SYNTH["cube3d_rk4_irs"]()
    q0 = 2.5e5
    s = gpu.Solver(c)
    _, vol, _ = s.metrics(1, 1)

    def hook(level, irk, block, T, R):
        R -= q0 * vol

    c2 = SYNTH["cube3d_rk4_irs"]()
    for b in c2.blocks:
        b.qvol[...] = b.qvol + q0
    o = oracle_mod.Oracle(c2)
    for _ in range(10):
        s.explicit_step(1, hook=hook)
        o.step(1)
    assert relerr(s.get("T"), o.get("T")) < RTOL
    s.close()
