"""GPU parity for the branches the closed-form oracle checks of
tests/test_oracle_pins3d.py pin: sheared 3-D blocks (tangential weights,
k faces), connections on skewed grids, symmetry (300) and radiative-
convective (305) walls, the state check's location, and the exact tile
shapes the benchmark runs on a block taller than one column."""
import numpy as np
import pytest

from fuss_amd import cases
from tests import cases3d as c3
from tests.test_gpu_parity import RTOL, relerr, steps_vs_oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gpu():
    from tests.conftest import gpu_available

    if not gpu_available():
        pytest.skip("no GPU")
    from fuss_amd import gpu as g

    g.load()
    return g


def lattice_walls(walls):
    return cases.lattice(nl=2, lo=4, hi=8, walls=walls)


PINS = {
    # quadratic field on the sheared affine lattice: every tangential weight
    # and every k-face term, 302/303/304/305 walls around it
    "affine_quadratic_walls": lambda: c3.affine_quadratic(
        n_rk=3, faces=[(302, cases.R(Tw=480.0)), (303, cases.R(hconv=40.0, Tref=300.0)),
                       (304, cases.R(eps=0.8, Tref=300.0)), (305, cases.R(eps=0.6, hconv=20.0, Tref=320.0)),
                       (301, cases.R(qw=2e3)), (302, cases.R(Tw=520.0))]),
    # the same lattice cut in two blocks along i, j and k (101 on skewed faces)
    "affine_split_i": lambda: c3.affine_split(0, n_rk=3),
    "affine_split_j": lambda: c3.affine_split(1, n_rk=3),
    "affine_split_k": lambda: c3.affine_split(2, n_rk=3, irs_beta=0.5),
    # symmetry faces: Cartesian box (3-D and 2-D), sheared block, and the
    # curvilinear connected lattice
    "box_300": lambda: c3.box(faces=[(300, cases.R()), (302, cases.R(Tw=520.0)), (300, cases.R()),
                                     (303, cases.R(hconv=40.0, Tref=300.0)), (300, cases.R()), (301, cases.R())]),
    "plate_300": lambda: c3.box((9, 7, 1), faces=[(300, cases.R()), (302, cases.R(Tw=520.0)), (300, cases.R()),
                                                  (301, cases.R(qw=-3e3)), (0, cases.R()), (0, cases.R())]),
    "affine_300": lambda: c3.affine_quadratic(
        n_rk=3, faces=[(300, cases.R()), (302, cases.R(Tw=480.0)), (300, cases.R()), (301, cases.R()),
                       (300, cases.R()), (303, cases.R(hconv=40.0, Tref=300.0))]),
    "lattice_300": lambda: lattice_walls([(300, cases.R()), (304, cases.R(eps=0.8, Tref=300.0)), (301, cases.R(qw=-5e3)),
                                          (300, cases.R()), (302, cases.R(Tw=550.0)), (300, cases.R())]),
    # 305 on the curvilinear lattice and the stretched box
    "lattice_305": lambda: lattice_walls([(302, cases.R(Tw=450.0)), (305, cases.R(eps=0.8, hconv=25.0, Tref=300.0)),
                                          (301, cases.R(qw=-5e3)), (305, cases.R(eps=0.5, hconv=60.0, Tref=350.0)),
                                          (302, cases.R(Tw=550.0)), (301, cases.R())]),
    "box_305": lambda: c3.box(faces=[(302, cases.R(Tw=600.0)), (305, cases.R(eps=0.7, hconv=35.0, Tref=300.0)),
                                     (301, cases.R()), (305, cases.R(eps=0.3, hconv=0.0, Tref=280.0)),
                                     (305, cases.R(eps=0.0, hconv=15.0, Tref=310.0)), (301, cases.R())], T0=520.0),
    # chimera (102) receivers on i, j and k faces of a 3-D block (cubic ghost
    # metric extrapolation across all three, the 3-D stage kernels and IRS
    # next to chimera ghosts), Cartesian and sheared; the sheared two-block
    # lattice joined by chimera entries along k
    "chimera3d": lambda: c3.chimera3d(),
    "chimera3d_sheared": lambda: c3.chimera3d(sheared=True),
    "chimera3d_irs": lambda: c3.chimera3d(irs_beta=0.5, n_rk=4),
    "chimera_split_k": lambda: c3.chimera_split(2, n_rk=3),
}


@pytest.mark.parametrize("name", list(PINS))
def test_pinned_branches_vs_oracle(gpu, oracle_mod, name):
    print(f"case {name}", flush=True)
    steps_vs_oracle(gpu, oracle_mod, PINS[name]())


def test_affine_quadratic_residual_on_device(gpu):
    """The device's first-stage update on the sheared lattice equals the
    closed form: interior cells change by k lap(T) dt / (rho cp) (R = -k V
    lap T)."""
    k = 10.0
    c = c3.affine_quadratic(k=k, n_rk=1)
    T0 = c.blocks[0].T.copy()
    s = gpu.Solver(c)
    s.explicit_step(1)
    dT = s.get("T") - T0
    dt = s.get("dt")
    expect = k * np.trace(c3.QUAD_H) * dt / (1000.0 * 500.0)
    inner = (slice(1, -1),) * 3
    np.testing.assert_allclose(dT[inner], expect[inner], rtol=1e-9)
    s.close()


@pytest.mark.parametrize("n_rk", [1, 3])
@pytest.mark.parametrize("graph", [False, True])
def test_state_check_location(gpu, oracle_mod, n_rk, graph):
    """Check_And_Fix_State: the step fails with FUSS_ESTATE and reports the
    first failing cell in the reference's loop order, like the oracle, both
    from fuss_step and from fuss_run's captured steps."""
    from tests.test_oracle_pins3d import oracle_step_status

    c = c3.drain_plate(cells=(5, 3), n_rk=n_rk)
    st_o, bijk_o = oracle_step_status(oracle_mod, oracle_mod.Oracle(c))
    assert st_o == gpu.FUSS_ESTATE and bijk_o == (1, 1, 3, 1)
    s = gpu.Solver(c)
    with pytest.raises(gpu.FussError) as ei:
        if graph:
            s.run_level(1, 60, 60, -1.0, 1e300, hist=False)
        else:
            s.explicit_step(1)
    assert ei.value.code == gpu.FUSS_ESTATE
    assert ei.value.bijk == bijk_o
    s.close()


def test_state_check_location_3d(gpu, oracle_mod):
    """The same on a 3-D box (stage kernels of 3-D levels): two drained
    cells on face 1 at (j, k) = (4, 3) and (2, 4); the reference stops at the
    first in (k, j, i) order, (1, 1, 4, 3)."""
    from tests.test_oracle_pins3d import oracle_step_status

    c = c3.box(n=(6, 5, 5), faces=c3.ADIABATIC, noise=0.0)
    ib, rb = c.bcs[0].ib, c.bcs[0].rb
    for j, k in ((2, 4), (4, 3)):
        rb[(ib[:, 4] == 1) & (ib[:, 2] == j) & (ib[:, 3] == k), 0] = 5e8
    st_o, bijk_o = oracle_step_status(oracle_mod, oracle_mod.Oracle(c))
    assert st_o == gpu.FUSS_ESTATE and bijk_o == (1, 1, 4, 3)
    s = gpu.Solver(c)
    with pytest.raises(gpu.FussError) as ei:
        s.explicit_step(1)
    assert ei.value.code == gpu.FUSS_ESTATE and ei.value.bijk == bijk_o
    s.close()


def test_benchmark_tiles_tall_cube(gpu, oracle_mod):
    """The benchmark's tile configuration (64-plane columns, 4-row LDS-ring
    stage kernel, 8-row fused x-y-z IRS pass) on a cube taller than one
    column whose k extent is not a multiple of 64 and whose i extent is not
    a multiple of the 64-cell tile row: RK4 + IRS steady primitive steps on
    a rectilinear level (the first stage computes dt), against the oracle."""
    c = cases.cube3d(n=8, mgl=1, n_rk=4)
    # replace the cube by a 72 x 40 x 136 stretched box with the same physics
    rng = np.random.Generator(np.random.PCG64(3))
    dims = (72, 40, 136)
    axes = [cases.stretch(m, 0.1 * m / 136, rng.uniform(0.2, 0.5), rng.uniform(0, 2 * np.pi)) for m in dims]
    nodes = cases.nodes_from_axes(*axes)
    T = 300.0 + 5.0 * rng.uniform(-1, 1, size=dims[::-1])
    from fuss_amd import case as cs

    blk = cs.Block(dims, nodes, T, np.ones_like(T), np.full_like(T, 1e6))
    faces = [(303, cases.R(hconv=50.0, Tref=300.0))] * 6
    bcs = [cases.bclist(*cases.single_block_bc(1, dims, faces))]
    c = cs.Case([blk], bcs, c.table, c.params, [], [np.full(bcs[0].n, -1, np.int32)])
    s = gpu.Solver(c)
    s.set_tiles(1, kc=64, row_march=0, stage_rows=4, irs_rows=8)
    assert s.tiles(1) == (64, 0, 4, 8)
    assert s.rect_info(1)[0] and s.rect_info(1)[1]
    o = oracle_mod.Oracle(c)
    for n in range(3):
        r = s.explicit_step(1)
        ro = o.step(1)
        assert abs(np.sqrt(r.res_sq) - np.sqrt(ro)) <= 1e-9 * np.sqrt(ro)
        assert relerr(s.get("T"), o.get("T")) < RTOL
    s.close()
