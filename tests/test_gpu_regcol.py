"""The register-column stage kernel (fuss_set_tiles stage_rows =
FUSS_STAGE_REGCOL: one wave per 64x4 column, no LDS ring, no barriers)
against the 4-row LDS-ring stage kernel and the oracle.

Both kernels evaluate every face flux, BC value and residual with the same
operations in the same order (Mod_Fluxes.f90:54-124, Lib_Newstate.f90:241-256),
so the state after any number of steps must agree bit for bit, and so must
dtglobal (the first stage's local dt on steady primitive levels).  The
residual's partial sums are grouped per column in both kernels but reduced
per wave here and per work-group there: equal to rounding.

Cases cover every variant of the kernel: steady primitive with IRS (the
first stage computes dt), conservative and primitive without IRS (the
update and the residual partials in the stage kernel), time-accurate steps
(global dt), blocks of mixed materials (per-cell material loads), a
non-uniform qvol, several blocks of different sizes, every wall type on a
rectilinear box, blocks wider than one 64-cell tile row and taller than a
column (edge faces, k0 > 1 prologues, ragged rows and columns)."""
import numpy as np
import pytest

from fuss_amd import case as cs
from fuss_amd import cases
from tests import cases3d as c3
from tests.test_gpu_parity import RTOL, SYNTH, relerr

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gpu():
    from tests.conftest import gpu_available

    if not gpu_available():
        pytest.skip("no GPU")
    from fuss_amd import gpu as g

    g.load()
    return g


WALLS = [(300, cases.R()), (301, cases.R(qw=4e3)), (302, cases.R(Tw=520.0)),
         (303, cases.R(hconv=40.0, Tref=300.0)), (304, cases.R(eps=0.8, Tref=300.0)),
         (305, cases.R(eps=0.6, hconv=20.0, Tref=320.0))]


def tallbox(dims=(72, 40, 44), faces=None, seed=3, n_rk=4, irs_beta=0.5, prim=True, time_accurate=False,
            vnn=2.0, nmat=1, qvol=1e6, qvol_noise=0.0, T0=300.0):
    """A stretched (rectilinear) box with a k(T) table, a noisy initial field,
    and optionally mixed materials per cell and a non-uniform qvol."""
    rng = np.random.Generator(np.random.PCG64(seed))
    axes = [cases.stretch(m, 0.1 * m / max(dims), rng.uniform(0.2, 0.5), rng.uniform(0, 2 * np.pi)) for m in dims]
    nodes = cases.nodes_from_axes(*axes)
    T = T0 + 5.0 * rng.uniform(-1, 1, size=dims[::-1])
    if nmat > 1:
        mat = rng.integers(1, nmat + 1, size=dims[::-1]).astype(np.float64)
        table = cases.material_tables(nmat=nmat, seed=seed)
        T = T + 200.0
    else:
        mat = np.ones_like(T)
        table = cases.cube_table(seed)
    qv = np.full_like(T, qvol) * (1.0 + qvol_noise * rng.uniform(-1, 1, size=T.shape))
    blk = cs.Block(dims, nodes, T, mat, qv)
    faces = faces or [(303, cases.R(hconv=50.0, Tref=300.0))] * 6
    bcs = [cases.bclist(*cases.single_block_bc(1, dims, faces))]
    p = cs.Params(n_rk=n_rk, vnn=vnn, prim=prim, irs_beta=irs_beta, time_accurate=time_accurate, mgl=1,
                  level_iter=[10**9], res_threshold=1e-10)
    return cs.Case([blk], bcs, table, p, [], [np.full(bcs[0].n, -1, np.int32)], name="tallbox")


CASES = {
    # RK4 + IRS steady primitive: dt in the first stage, 2 tile rows in i,
    # columns of 8/16/32 planes over nk = 44 (ragged last column)
    "tall_rk4_irs": lambda: tallbox(),
    # every wall type on the six faces (BC values on x, y and z tiles)
    "walls_rk3_irs": lambda: tallbox((70, 9, 19), faces=WALLS, n_rk=3, T0=520.0),
    # no IRS: the stage kernel applies the update (conservative and primitive)
    "cons_noirs": lambda: tallbox((66, 10, 21), n_rk=3, irs_beta=0.0, prim=False, vnn=0.8),
    "prim_noirs": lambda: tallbox((66, 10, 21), n_rk=3, irs_beta=0.0, vnn=0.8),
    # time-accurate (global dt), with and without IRS
    "transient_irs": lambda: tallbox((65, 12, 17), n_rk=3, time_accurate=True, vnn=0.8),
    "transient_cons": lambda: tallbox((65, 12, 17), n_rk=4, irs_beta=0.0, prim=False, time_accurate=True, vnn=0.8),
    # mixed materials per cell (per-cell material and table row bases)
    "multimat_irs": lambda: tallbox((67, 11, 18), nmat=3, n_rk=3),
    "multimat_cons": lambda: tallbox((67, 11, 18), nmat=3, n_rk=4, irs_beta=0.0, prim=False, vnn=0.5),
    # a non-uniform qvol (the qvol stream)
    "qvol_field": lambda: tallbox((64, 8, 16), qvol_noise=0.5),
    # several rectilinear blocks of different sizes and materials
    "sweep3_rk4_irs": SYNTH["sweep3_rk4_irs"],
    # the 24^3 cube (one partial tile row) on the finest of three levels
    "cube3d_mg3_irs": SYNTH["cube3d_mg3_irs"],
}


def run(gpu, c, kc, stage_rows, irs_rows=4, steps=5, precision=None):
    s = gpu.Solver(c) if precision is None else gpu.Solver(c, precision=precision)
    s.set_tiles(1, kc=kc, row_march=0, stage_rows=stage_rows, irs_rows=irs_rows)
    assert s.tiles(1) == (kc, 0, stage_rows, irs_rows)
    rs = [s.explicit_step(1) for _ in range(steps)]
    T = np.concatenate([x.ravel() for x in s.T()])
    s.close()
    return T, np.array([r.res_sq for r in rs]), np.array([r.dtglobal for r in rs])


@pytest.mark.parametrize("kc", [8, 16, 32])
@pytest.mark.parametrize("name", list(CASES))
def test_regcol_equals_ring(gpu, name, kc):
    c = CASES[name]()
    ring = run(gpu, c, kc, 4)
    reg = run(gpu, c, kc, gpu.FUSS_STAGE_REGCOL)
    assert np.array_equal(reg[0], ring[0])
    np.testing.assert_allclose(reg[1], ring[1], rtol=1e-13)
    assert np.array_equal(reg[2], ring[2])


@pytest.mark.parametrize("name", ["tall_rk4_irs", "walls_rk3_irs", "multimat_cons", "transient_irs"])
def test_regcol_vs_oracle(gpu, oracle_mod, name):
    """Three steps with register columns (and the 8-row IRS pass where the
    level takes it) against the oracle."""
    c = CASES[name]()
    s = gpu.Solver(c)
    s.set_tiles(1, kc=16, row_march=0, stage_rows=gpu.FUSS_STAGE_REGCOL, irs_rows=8)
    o = oracle_mod.Oracle(c)
    for _ in range(3):
        r = s.explicit_step(1)
        ro = o.step(1)
        assert abs(np.sqrt(r.res_sq) - np.sqrt(ro)) <= 1e-9 * np.sqrt(ro)
        assert relerr(s.get("T"), o.get("T")) < RTOL
    s.close()


def test_regcol_graph_run(gpu):
    """fuss_run (captured graphs) with register columns gives the step loop's
    state and residual history bit for bit."""
    out = []
    for graphs in (True, False):
        s = gpu.Solver(CASES["tall_rk4_irs"]())
        s.set_tiles(1, kc=16, row_march=0, stage_rows=gpu.FUSS_STAGE_REGCOL, irs_rows=8)
        s.graphs(graphs)
        k, hr, _, _ = s.run_level(1, 36, 36, -1.0, 1e300)
        out.append((k, hr.copy(), np.concatenate([T.ravel() for T in s.T()])))
        s.close()
    assert out[0][0] == out[1][0] == 36
    assert np.array_equal(out[0][1], out[1][1]) and np.array_equal(out[0][2], out[1][2])


def test_regcol_refused_where_it_cannot_run(gpu):
    """fuss_set_tiles refuses register columns on curvilinear levels, at
    column heights it is not built for, and with the per-axis factors off."""
    s = gpu.Solver(SYNTH["lattice_transient"]())
    with pytest.raises(gpu.FussError) as ei:
        s.set_tiles(1, kc=16, row_march=0, stage_rows=gpu.FUSS_STAGE_REGCOL)
    assert ei.value.code == gpu.FUSS_EUNSUPPORTED
    s.close()
    s = gpu.Solver(CASES["qvol_field"]())
    for kc in (4, 64):
        with pytest.raises(gpu.FussError):
            s.set_tiles(1, kc=kc, row_march=0, stage_rows=gpu.FUSS_STAGE_REGCOL)
    s.rect(False)
    with pytest.raises(gpu.FussError):
        s.set_tiles(1, kc=16, row_march=0, stage_rows=gpu.FUSS_STAGE_REGCOL)
    s.close()


def test_regcol_table_error(gpu):
    """A cell beyond the property table stops the step with FUSS_ETABLE."""
    c = tallbox((66, 9, 12), n_rk=3)
    c.blocks[0].T[5, 4, 40] = 2.0e4
    s = gpu.Solver(c)
    s.set_tiles(1, kc=8, row_march=0, stage_rows=gpu.FUSS_STAGE_REGCOL)
    with pytest.raises(gpu.FussError) as ei:
        s.explicit_step(1)
    assert ei.value.code == gpu.FUSS_ETABLE
    s.close()


def _graded(dims, axis, lo, hi, seed=5):
    """A tall box whose temperature runs from lo to hi along one axis."""
    c = tallbox(dims, seed=seed)
    T = c.blocks[0].T
    n = T.shape[2 - axis]
    ramp = lo + (hi - lo) * np.linspace(0.0, 1.0, n)
    shape = [1, 1, 1]
    shape[2 - axis] = n
    c.blocks[0].T = T - 300.0 + ramp.reshape(shape)
    return c


# Steady primitive stages with IRS on a level of one material stage the
# table rows of each column in LDS (a window of rows around its first
# plane's temperatures); a face or cell outside the window makes the
# work-group compute its column again from memory, and a first plane whose
# rows span half the window sends the column to memory at once.
WINDOW_CASES = {
    # 300 -> 700 K along k: every column leaves its window partway up
    "ramp_k": lambda: _graded((66, 9, 70), 2, 300.0, 700.0),
    # 300 -> 900 K along i: first planes span more than half the window
    "ramp_i": lambda: _graded((70, 9, 20), 0, 300.0, 900.0),
    # one hot cell in one column (that work-group alone misses)
    "hot_cell": lambda: _hot_cell(),
}


def _hot_cell():
    c = tallbox((70, 12, 40))
    c.blocks[0].T[30, 6, 20] = 820.0
    return c


@pytest.mark.parametrize("fp32", [False, True])
@pytest.mark.parametrize("stage_rows", [4, 8])
@pytest.mark.parametrize("name", list(WINDOW_CASES))
def test_window_misses_match_memory_path(gpu, name, stage_rows, fp32):
    """The LDS-ring kernel with its row window (and the recomputed columns)
    gives the register-column kernel's state (rows from memory) bit for
    bit, in FP64 and in FP32 contexts."""
    c = WINDOW_CASES[name]()
    prec = gpu.FUSS_FP32 if fp32 else None
    ring = run(gpu, c, 16, stage_rows, precision=prec)
    reg = run(gpu, c, 16, gpu.FUSS_STAGE_REGCOL, precision=prec)
    assert np.array_equal(ring[0], reg[0])
    np.testing.assert_allclose(ring[1], reg[1], rtol=1e-13)
    assert np.array_equal(ring[2], reg[2])


@pytest.mark.parametrize("name", list(WINDOW_CASES))
def test_window_misses_vs_oracle(gpu, oracle_mod, name):
    c = WINDOW_CASES[name]()
    s = gpu.Solver(c)
    s.set_tiles(1, kc=64, row_march=0, stage_rows=8, irs_rows=8)
    o = oracle_mod.Oracle(c)
    for _ in range(3):
        r = s.explicit_step(1)
        ro = o.step(1)
        assert abs(np.sqrt(r.res_sq) - np.sqrt(ro)) <= 1e-9 * np.sqrt(ro)
        assert relerr(s.get("T"), o.get("T")) < RTOL
    s.close()


def _short_table_box():
    """A table of 100 rows (250-349 K), shorter than the row window: the
    ring kernel reads its rows from memory."""
    from fuss_amd import formats as fm

    c = tallbox((66, 9, 20))
    Tr = np.arange(250, 350, dtype=np.float64)
    kap = (20.0 + 0.02 * (Tr - 300.0)).reshape(1, -1)
    rho = np.full_like(kap, 7800.0)
    cp = (460.0 + 0.1 * (Tr - 300.0)).reshape(1, -1)
    c.table = fm.PropertyTable(250, cp, rho, kap, np.cumsum(rho * cp, axis=1), ["short"])
    return c


@pytest.mark.parametrize("stage_rows", [4, 8])
def test_short_table_memory_path(gpu, oracle_mod, stage_rows):
    c = _short_table_box()
    ring = run(gpu, c, 16, stage_rows)
    reg = run(gpu, c, 16, gpu.FUSS_STAGE_REGCOL)
    assert np.array_equal(ring[0], reg[0])
    assert np.array_equal(ring[2], reg[2])
    s = gpu.Solver(c)
    s.set_tiles(1, kc=16, row_march=0, stage_rows=stage_rows, irs_rows=8)
    o = oracle_mod.Oracle(c)
    for _ in range(2):
        s.explicit_step(1)
        o.step(1)
    assert relerr(s.get("T"), o.get("T")) < RTOL
    s.close()
