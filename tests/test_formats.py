"""Round trips of the reference file formats (fuss_amd.formats) and the
case-directory writer: a case written as input.ini + INPUT/ reads back
identically, and a reference case directory loads with the reference's BC
entry count."""
import os

import numpy as np

from fuss_amd import case as cs
from fuss_amd import cases
from fuss_amd import formats as fm


def test_case_dir_round_trip(tmp_path):
    c = cases.lattice(nl=2, lo=3, hi=5, seed=11, mgl=1)
    cases.write_case_dir(c, str(tmp_path))
    d = cs.load_case_dir(str(tmp_path))
    assert d.nb == c.nb
    for a, b in zip(c.blocks, d.blocks):
        assert a.dims == b.dims
        np.testing.assert_array_equal(a.nodes, b.nodes)
        np.testing.assert_array_equal(a.T, b.T)
        np.testing.assert_array_equal(a.matID, b.matID)
        np.testing.assert_array_equal(a.qvol, b.qvol)
    np.testing.assert_array_equal(c.bcs[0].ib, d.bcs[0].ib)
    np.testing.assert_array_equal(c.bcs[0].rb, d.bcs[0].rb)
    for f in ("cp", "rho", "kap", "h"):
        np.testing.assert_array_equal(getattr(c.table, f), getattr(d.table, f))
    assert d.params.n_rk == c.params.n_rk and d.params.vnn == c.params.vnn
    assert d.params.time_accurate == c.params.time_accurate and d.params.prim == c.params.prim


def test_bc_types_round_trip(tmp_path):
    ib = [[1, 1, 1, 1, 1, t] + ([2, 1, 1, 1, 2, 1, 0, 0, 1] if t == 101 else [0] * 9)
          for t in (0, 101, 300, 301, 302, 303, 304, 305)]
    rb = [[1.5, 2.5, 3.5, 4.5, 5.5, 0.75]] * len(ib)
    bcl = cases.bclist(ib, rb)
    p = os.path.join(str(tmp_path), "bc.txt")
    fm.write_bc(p, bcl)
    r = fm.read_bc(p)
    np.testing.assert_array_equal(r.ib, bcl.ib)
    # each type keeps the reals it carries (IO_BC.f90:173-216)
    keep = {0: [], 101: [], 300: [], 301: [0], 302: [1], 303: [2, 3, 4], 304: [5, 4], 305: [2, 5, 4]}
    for n, t in enumerate(r.ib[:, 5]):
        for q in range(6):
            expect = bcl.rb[n, q] if q in keep[int(t)] else 0.0
            assert r.rb[n, q] == expect


def test_chimera_bc_round_trip(tmp_path):
    """102 entries carry their donor lists (IO_BC.f90:218-229): the chimera
    fixture's BC list (the reference's own bc.txt as read) writes and reads
    back with the same entries, layer counts, donors and volume fractions,
    and in the layout fuss_set_chimera takes."""
    from tests.golden import fixtures

    c, _ = fixtures.load("chimera")
    bcl = c.bcs[0]
    assert len(bcl.chim) == 17 and all(int(t) == 102 for t in bcl.ib[sorted(bcl.chim), 5])
    p = os.path.join(str(tmp_path), "bc.txt")
    fm.write_bc(p, bcl)
    r = fm.read_bc(p)
    np.testing.assert_array_equal(r.ib, bcl.ib)
    assert sorted(r.chim) == sorted(bcl.chim)
    for e, (n1, don, vf) in bcl.chim.items():
        assert r.chim[e][0] == n1
        np.testing.assert_array_equal(r.chim[e][1], don)
        np.testing.assert_array_equal(r.chim[e][2], vf)
    ent, n1, n2, don, vf = r.chimera_arrays()
    assert len(ent) == 17 and don.shape == (int(n1.sum() + n2.sum()), 4) and vf.shape == (don.shape[0],)
    # the layer-1 fractions of each ghost sum to one (a partition of its volume)
    off = 0
    for a, b in zip(n1, n2):
        assert abs(vf[off:off + a].sum() - 1.0) < 1e-8
        off += a + b


def test_bc_count_check():
    c = cases.plate2d(n=8)
    fm.check_bc_count(c.bcs[0], [b.dims for b in c.blocks])
    bad = fm.BCList(c.bcs[0].ib[:-1], c.bcs[0].rb[:-1], [None] * (c.bcs[0].n - 1))
    try:
        fm.check_bc_count(bad, [b.dims for b in c.blocks])
    except ValueError as e:
        assert "different" in str(e)
    else:
        raise AssertionError("truncated BC list accepted")


def test_series_and_ini(tmp_path):
    p = os.path.join(str(tmp_path), "s.dat")
    with open(p, "w") as fh:
        fh.write("periodic\n0.0 1.0\n1.0 3.0\n2.0 1.0\n")
    per, t, v = fm.read_series(p)
    assert per and list(t) == [0.0, 1.0, 2.0] and list(v) == [1.0, 3.0, 1.0]
    q = os.path.join(str(tmp_path), "input.ini")
    with open(q, "w") as fh:
        fh.write("[FUSS-Numerics]\ntime-scheme = RK2\nirs-beta = 0.5\n;\n[FUSS-Multigrid]\nlevels = 2\nlevel1-iter = 7\n")
    prm = cs.params_from_ini(fm.read_ini(q))
    assert prm.n_rk == 2 and prm.irs_beta == 0.5 and prm.mgl == 2 and prm.itermax() == [7, 0]
    assert not prm.prim  # cons is the reference default (Register_Numerics.f90:40)
