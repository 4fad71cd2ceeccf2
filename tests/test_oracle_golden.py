"""Pin the CPU oracle to the reference's own stored outputs.

Each fixture (tests/golden/*.npz, written by make_fixtures.py from the
reference's test/ directory) runs through the oracle with the reference's
inputs and numerics; the oracle must reproduce the stored solution:

* steady cases: the stop iteration (first residual <= res-threshold), the
  final T field (Tecplot output, 15 significant digits) and the wall output
  (wall.tec: Tw, qw of every wall face cell) -- among them the chimera case
  (two blocks of different j resolution joined by 102 entries: ghost
  enthalpy from volume-fraction-weighted donors, 90 204 iterations);
* multimat_Plate: six transient snapshots (time and T);
* oscillating_Rod: both probe histories (time, T) of a periodic 302 series;
* mg_irs: all 80603 residuals of a 2-level multigrid + IRS + connection run.
"""
import numpy as np
import pytest

from fuss_amd import formats as fm
from tests import known_answers
from tests.golden import fixtures

STEADY = ["temperature_Plate", "qvol_Plate", "convective_Plate", "radiative_Rod", "chimera"]
# Tecplot stores 15 significant digits; the oracle follows the reference's
# operation order, so agreement is at that level.
FIELD_RTOL = 1e-13
# wall.tec is printed the same way; qw (a conductance times a difference of
# nearby temperatures, scaled by the zone's largest |qw|) keeps a few digits
# fewer: observed <= 6e-14
WALL_RTOL = 1e-12


@pytest.mark.parametrize("name", STEADY)
def test_steady_field(oracle_mod, name):
    c, z = fixtures.load(name)
    stop = int(z["stop_iter"])
    o = oracle_mod.Oracle(c)
    n, hist, _ = o.run(max_iters=stop + 5)
    assert n == stop, f"oracle stopped at {n}, reference at {stop}"
    assert hist[-1, 2] <= c.params.res_threshold < hist[-2, 2]
    for b in range(c.nb):
        ref = z[f"ref_T{b}"]
        err = np.max(np.abs(o.get("T", 1, b + 1) - ref) / np.abs(ref))
        assert err < FIELD_RTOL, f"block {b+1}: max rel err {err:.3e}"
    Tw, qw = o.wall(1)
    check_wall(c, z, Tw, qw)
    # the published known answer the reference's verify.py checks (< 1 %)
    known_answers.check_steady(c, name, [o.get("T", 1, b + 1) for b in range(c.nb)], Tw)


def check_wall(c, z, Tw, qw):
    """Wall output (Extract_Properties_From_BC) of the converged state against
    the reference's wall.tec: the same wall zones, Tw and qw per face cell.
    qw is scaled by the zone's largest |qw| (it crosses 0 on some walls)."""
    zones = fm.wall_zones(c.bcs[0], [b.dims for b in c.blocks], Tw, qw)
    names = {k[5:-3] for k in z.files if k.startswith("wall_") and k.endswith("_Tw")}
    assert {f"B{b}F{f}" for b, f in zones} == names
    for (b, f), v in zones.items():
        rT, rq = z[f"wall_B{b}F{f}_Tw"], z[f"wall_B{b}F{f}_qw"]
        assert v["Tw"].shape == rT.shape
        assert np.max(np.abs(v["Tw"] - rT) / np.abs(rT)) < WALL_RTOL, (b, f)
        assert np.max(np.abs(v["qw"] - rq)) <= WALL_RTOL * max(np.max(np.abs(rq)), 1e-300), (b, f)


def test_multimat_transient(oracle_mod):
    c, z = fixtures.load("multimat_Plate")
    o = oracle_mod.Oracle(c)
    it = 0
    for q, target in enumerate(z["snap_iter"]):
        while it < target:
            o.step(1)
            it += 1
        assert abs(o.time(1) - float(z[f"snap{q}_time"])) <= 1e-14 * abs(float(z[f"snap{q}_time"])) + 1e-16
        for b in range(c.nb):
            ref = z[f"snap{q}_T{b}"]
            err = np.max(np.abs(o.get("T", 1, b + 1) - ref) / np.abs(ref))
            assert err < FIELD_RTOL, f"snapshot {q} block {b+1}: {err:.3e}"
        # against the slab's Fourier series, as the reference's verify.py (< 0.02)
        assert known_answers.multimat_error(c, o.get("T", 1, 2), o.time(1)) < 0.02


def test_oscillating_probes(oracle_mod):
    c, z = fixtures.load("oscillating_Rod")
    o = oracle_mod.Oracle(c)
    n, hist, pr = o.run(max_iters=10**6, probes=z["probes"])
    p1, p2 = z["p1"], z["p2"]
    k = p1.shape[0]
    assert pr.shape[0] >= k
    np.testing.assert_allclose(pr[:k, 0], p1[:, 0], rtol=5e-14)
    np.testing.assert_allclose(pr[:k, 1], p1[:, 1], rtol=FIELD_RTOL)
    np.testing.assert_allclose(pr[:k, 2], p2[:, 1], rtol=FIELD_RTOL)
    # the NAFEMS peak temperature rise (< 1 %)
    assert known_answers.oscillating_error(pr[:k, 1], pr[:k, 2]) < 0.01


def residual_tolerance(res, Tmax, ncell):
    """The stored residual has 10 significant digits; late in the run the
    residual sqrt(sum dT^2) is itself rounding-limited: every cell's dT is
    only known to ~ulp(T), so allow sqrt(ncell)*ulp(Tmax) on top."""
    return 6e-10 * res + 4.0 * np.sqrt(ncell) * np.spacing(Tmax)


@pytest.mark.slow
def test_mg_irs_trajectory(oracle_mod):
    c, z = fixtures.load("mg_irs")
    ref = z["res_hist"]
    o = oracle_mod.Oracle(c)
    n, hist, _ = o.run(max_iters=ref.size + 10)
    assert n == ref.size, f"oracle took {n} iterations, reference {ref.size}"
    np.testing.assert_array_equal(hist[:, 0], np.arange(1, n + 1))
    tol = residual_tolerance(ref, 2000.0, c.ncell)
    bad = np.abs(hist[:, 2] - ref) > tol
    assert not bad.any(), f"{bad.sum()} residuals off, first at iteration {np.argmax(bad)+1}"
    # the first 10000 iterations agree to the stored 10 digits
    np.testing.assert_allclose(hist[:10000, 2], ref[:10000], rtol=6e-10)
