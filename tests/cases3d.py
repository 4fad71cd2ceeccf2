"""Small 3-D single-block cases for the closed-form checks of the oracle and
their GPU parity runs (built in memory in the reference's input layout)."""
import numpy as np

from fuss_amd import case as cs
from fuss_amd import cases

# a shear with every off-diagonal term nonzero: the cells are identical
# parallelepipeds whose edges are neither orthogonal nor axis-aligned
SHEAR = np.array([[1.0, 0.30, 0.10],
                  [0.20, 1.0, 0.25],
                  [0.15, -0.10, 1.0]])


def affine_nodes(n, h, S=SHEAR, x0=(0.02, 0.01, 0.03)):
    """Nodes (nk+1, nj+1, ni+1, 3) of the lattice X = S (i h0, j h1, k h2) + x0."""
    i, j, k = (np.arange(m + 1) * hh for m, hh in zip(n, h))
    K, J, I = np.meshgrid(k, j, i, indexing="ij")
    E = np.stack([I, J, K], axis=-1)
    return np.ascontiguousarray(E @ np.asarray(S).T + np.asarray(x0))


def cell_centres(nodes):
    """Centroids of the cells of an affine lattice (mean of the 8 nodes)."""
    c = 0.0
    for a in (0, 1):
        for b in (0, 1):
            for d in (0, 1):
                c = c + nodes[a:a + nodes.shape[0] - 1, b:b + nodes.shape[1] - 1, d:d + nodes.shape[2] - 1]
    return c / 8.0


# T = T0 + g.x + x.Hx / 2 with every second-derivative term nonzero
QUAD_G = np.array([50.0, -30.0, 20.0])
QUAD_H = 1000.0 * np.array([[2.0, 0.3, 0.2],
                            [0.3, 1.0, -0.4],
                            [0.2, -0.4, 1.6]])


def quad_field(X, T0=500.0, g=QUAD_G, H=QUAD_H):
    return T0 + X @ g + 0.5 * np.einsum("...i,ij,...j->...", X, H, X)


def single_block(nodes, T, faces, k=10.0, rho=1000.0, cp=500.0, n_rk=3, vnn=0.5, prim=True, irs_beta=0.0,
                 time_accurate=False, qvol=0.0, table=None):
    nk, nj, ni = T.shape
    blk = cs.Block((ni, nj, nk), np.ascontiguousarray(nodes), np.ascontiguousarray(T, dtype=np.float64),
                   np.ones_like(T), np.full_like(T, qvol))
    bcs = [cases.bclist(*cases.single_block_bc(1, (ni, nj, nk), faces))]
    p = cs.Params(n_rk=n_rk, vnn=vnn, prim=prim, irs_beta=irs_beta, time_accurate=time_accurate, mgl=1,
                  level_iter=[10**9], res_threshold=1e-10)
    tab = table if table is not None else cases.const_table(k, rho, cp)
    return cs.Case([blk], bcs, tab, p, [], [np.full(bcs[0].n, -1, np.int32)])


ADIABATIC = [(301, cases.R())] * 6


def affine_quadratic(n=(8, 7, 6), h=(0.01, 0.012, 0.009), k=10.0, n_rk=1, faces=None, **kw):
    """Quadratic temperature field on a sheared affine 3-D lattice, constant k."""
    nodes = affine_nodes(n, h)
    T = quad_field(cell_centres(nodes))
    return single_block(nodes, T, faces or ADIABATIC, k=k, n_rk=n_rk, **kw)


def stretched_axes(n, seed=13, L=0.05):
    rng = np.random.default_rng(seed)
    return [np.concatenate([[0.0], np.cumsum(rng.uniform(0.6, 1.4, m))]) * (L / m) for m in n]


def zslab(n=(6, 5, 10), Tlo=350.0, Thi=650.0, faces_xy=None, seed=13):
    """Rectilinear box (stretched in x and y, uniform in z: the face metric is
    the mean of the two cells' M, exact for a linear field only at uniform
    spacing) with T linear in z between two 302 walls (faces 5, 6) whose
    temperatures continue the field to the wall planes; the other faces
    adiabatic (301, q = 0) unless faces_xy overrides faces 1-4."""
    x, y, _ = stretched_axes(n, seed)
    z = np.linspace(0.0, 0.04, n[2] + 1)
    nodes = cases.nodes_from_axes(x, y, z)
    zc = 0.5 * (z[1:] + z[:-1])
    slope = (Thi - Tlo) / (z[-1] - z[0])
    T = np.broadcast_to((Tlo + slope * (zc - z[0]))[:, None, None], (n[2], n[1], n[0])).copy()
    faces = (faces_xy or [(301, cases.R())] * 4) + [(302, cases.R(Tw=Tlo)), (302, cases.R(Tw=Thi))]
    return single_block(nodes, T, faces)


def box(n=(7, 6, 5), faces=None, T0=420.0, seed=13, noise=20.0, **kw):
    """Rectilinear (stretched, exactly Cartesian) box with a seeded non-uniform
    initial field and the given six (type, rrec) faces."""
    x, y, z = stretched_axes(n, seed)
    nodes = cases.nodes_from_axes(x, y, z)
    rng = np.random.default_rng(seed + 1)
    T = T0 + noise * rng.uniform(-1.0, 1.0, size=(n[2], n[1], n[0]))
    return single_block(nodes, T, faces or ADIABATIC, **kw)


def drain_plate(qw=2e6, cells=(3, 5), n_rk=3):
    """2-D plate whose face-1 entries at rows j in `cells` carry a 301 flux
    large enough to drive those cells below 0 K in the first stage (the
    other walls adiabatic)."""
    from tests.test_oracle_closed_forms import plate

    c = plate(n=8, m=6, dx=0.1, dy=0.05, faces=[(301, cases.R())] * 4 + [(0, cases.R())] * 2, n_rk=n_rk,
              vnn=0.5)
    ib, rb = c.bcs[0].ib, c.bcs[0].rb
    for j in cells:
        e = np.nonzero((ib[:, 4] == 1) & (ib[:, 2] == j))[0]
        rb[e, 0] = qw
    return c


def affine_split(axis, n=(8, 7, 6), h=(0.01, 0.012, 0.009), T_of_X=None, k=10.0, n_rk=1, walls=None,
                 m=None, **kw):
    """The sheared affine lattice cut in two blocks along `axis` (0, 1, 2) at
    cell m (default n/2), joined 1-to-1 by 101 entries on the cut faces
    (donor stencils d11..d22 = 1, 0, 0, 1 as the reference's lattice input
    writes them); the outer faces are `walls` (default adiabatic 301)."""
    nodes = affine_nodes(n, h)
    X = cell_centres(nodes)
    T = T_of_X(X) if T_of_X is not None else quad_field(X)
    m = m or n[axis] // 2
    ax = 2 - axis  # array axis of `axis` in (k, j, i) order
    cut = [slice(None)] * 3
    parts = []
    for lo, hi in ((0, m), (m, n[axis])):
        sn = list(cut)
        sn[ax] = slice(lo, hi + 1)
        sc = list(cut)
        sc[ax] = slice(lo, hi)
        Tb = np.ascontiguousarray(T[tuple(sc)])
        dims = (Tb.shape[2], Tb.shape[1], Tb.shape[0])
        parts.append((dims, np.ascontiguousarray(nodes[tuple(sn)]), Tb))
    walls = walls or ADIABATIC
    fhi, flo = 2 * axis + 2, 2 * axis + 1
    ib, rb, blocks = [], [], []
    for b, (dims, nd, Tb) in enumerate(parts):
        blocks.append(cs.Block(dims, nd, Tb, np.ones_like(Tb), np.zeros_like(Tb)))
        for f in range(1, 7):
            if (b == 0 and f == fhi) or (b == 1 and f == flo):
                other = parts[1 - b][0]
                fs = flo if b == 0 else fhi
                don = 1 if b == 0 else other[axis]

                def extra(i, j, kk, other_b=2 - b, fs=fs, don=don):
                    q = [i.copy(), j.copy(), kk.copy()]
                    q[axis][:] = don
                    return np.stack([np.full_like(i, other_b), q[0], q[1], q[2], np.full_like(i, fs),
                                     np.ones_like(i), np.zeros_like(i), np.zeros_like(i), np.ones_like(i)], axis=1)

                x, y = cases.face_entries(b + 1, dims, f, 101, cases.R(), extra)
            else:
                x, y = cases.face_entries(b + 1, dims, f, *walls[f - 1])
            ib.append(x)
            rb.append(y)
    bcs = [cases.bclist(np.concatenate(ib), np.concatenate(rb))]
    p = cs.Params(n_rk=n_rk, vnn=kw.pop("vnn", 0.5), prim=True, irs_beta=kw.pop("irs_beta", 0.0),
                  time_accurate=kw.pop("time_accurate", False), mgl=1, level_iter=[10**9], res_threshold=1e-10)
    return cs.Case(blocks, bcs, cases.const_table(k, 1000.0, 500.0), p, [], [np.full(bcs[0].n, -1, np.int32)])


def chimera3d(n_rk=3, irs_beta=0.0, sheared=False):
    """Two 3-D blocks side by side in x with different j / k counts: A (6 x 5 x
    4 cells) receives chimera (102) ghosts on its faces 2 (i), 4 (j) and 6
    (k) from donors among the cells of B (7 x 6 x 5); the other faces are
    walls (302, 303, 301).  Per entry two first-layer donors (volume fractions
    0.6, 0.4) and one second-layer donor, all interior cells of B.  The donor
    choice is arbitrary (a parity case, not a physical overlap).  sheared:
    both blocks on a sheared lattice (tangential weights)."""
    dims = [(6, 5, 4), (7, 6, 5)]
    h = (0.01, 0.012, 0.009)
    blocks, ib, rb = [], [], []
    chim = {}
    rng = np.random.default_rng(21)
    for b, (n, x0) in enumerate(zip(dims, [(0.0, 0.0, 0.0), (0.07, 0.0, 0.0)])):
        S = SHEAR if sheared else np.eye(3)
        nodes = affine_nodes(n, h, S=S, x0=x0)
        T = 400.0 + 40.0 * rng.uniform(-1, 1, size=n[::-1])
        blocks.append(cs.Block(n, nodes, T, np.ones_like(T), np.zeros_like(T)))
    walls = [(302, cases.R(Tw=450.0)), (303, cases.R(hconv=30.0, Tref=300.0)), (301, cases.R(qw=-2e3)),
             (302, cases.R(Tw=380.0)), (303, cases.R(hconv=60.0, Tref=320.0)), (301, cases.R())]
    nB = dims[1]
    for b, n in enumerate(dims):
        for f in range(1, 7):
            if b == 0 and f in (2, 4, 6):
                x, y = cases.face_entries(1, n, f, 102, cases.R())
                for r in range(x.shape[0]):
                    i, j, k = (int(v) for v in x[r, 1:4])
                    if f == 2:
                        d = [(1, min(j, nB[1]), min(k, nB[2])), (2, min(j, nB[1]), min(k, nB[2])),
                             (3, min(j, nB[1]), min(k, nB[2]))]
                    elif f == 4:
                        d = [(min(i, nB[0]), 1, min(k, nB[2])), (min(i, nB[0]), 2, min(k, nB[2])),
                             (min(i, nB[0]), 3, min(k, nB[2]))]
                    else:
                        d = [(min(i, nB[0]), min(j, nB[1]), 1), (min(i, nB[0]), min(j, nB[1]), 2),
                             (min(i, nB[0]), min(j, nB[1]), 3)]
                    don = np.array([(2,) + q for q in d], np.int32)
                    chim[sum(a.shape[0] for a in ib) + r] = (2, don, np.array([0.6, 0.4, 1.0]))
            else:
                x, y = cases.face_entries(b + 1, n, f, *walls[f - 1])
            ib.append(x)
            rb.append(y)
    bcl = cases.bclist(np.concatenate(ib), np.concatenate(rb))
    bcl.chim = chim
    p = cs.Params(n_rk=n_rk, vnn=0.5, prim=True, irs_beta=irs_beta, mgl=1, level_iter=[10**9], res_threshold=1e-10)
    return cs.Case(blocks, [bcl], cases.const_table(15.0, 3000.0, 600.0), p, [], [np.full(bcl.n, -1, np.int32)])


def chimera_split(axis, T_of_X=None, k=10.0, n_rk=1, **kw):
    """affine_split's two blocks with the cut faces joined by chimera (102)
    entries instead of 101: each entry's first-layer donor is the cell of the
    other block that its ghost cell coincides with (volume fraction 1), its
    second-layer donor the next cell inward."""
    c = affine_split(axis, T_of_X=T_of_X, k=k, n_rk=n_rk, **kw)
    bcl = c.bcs[0]
    ib = bcl.ib.copy()
    chim = {}
    for e in np.nonzero(ib[:, 5] == 101)[0]:
        bs, i_s, js, ks, fs = (int(v) for v in ib[e, 6:11])
        second = [i_s, js, ks]
        second[axis] += 1 if fs % 2 == 1 else -1
        don = np.array([[bs, i_s, js, ks], [bs] + second], np.int32)
        chim[int(e)] = (1, don, np.array([1.0, 1.0]))
        ib[e, 5] = 102
    nb = cases.bclist(ib, bcl.rb.copy())
    nb.chim = chim
    c.bcs[0] = nb
    return c
