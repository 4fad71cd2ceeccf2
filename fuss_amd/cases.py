"""Synthetic cases of the benchmark configurations, built in memory in the
reference's input layout (and writable as a reference case directory).

* ``plate2d``  -- single-block steady plate (uniform Cartesian, 1 cell in z):
  fixed-T faces 1/2, adiabatic 301 q=0 faces 3/4, null faces 5/6, constant
  properties, seeded uniform IC noise.  128x128, k=50, rho=7800, cp=460,
  T=500/300 is the 2-D parity configuration.
* ``cube3d``   -- single-block steady cube with seeded smooth per-axis
  stretching, a k(T) table from 32 noisy knots resampled to 1 K rows,
  uniform qvol and convective 303 walls; 256^3 with IRS and 3 levels is the
  3-D throughput configuration.
* ``lattice``  -- multi-block 4x4x4 lattice of curvilinear blocks joined by
  1-to-1 connections (101), three tabulated materials, mixed 302/301/304 walls.
* ``plate_adjoint`` -- 2-D transient plate with a segmented bottom heat flux.
* ``sweep``    -- independent single-block steady problems as one multi-block
  domain (no connections).

Tables follow the reference's properties.dat convention: one row per kelvin
from T=1, columns cp, rho, k, h with h(T) = sum_{t<=T} rho(t) cp(t) (e.g.
test/steady-state/temperature_Plate/INPUT/properties.dat: h(1)=rho*cp).
"""
from __future__ import annotations

import os

import numpy as np

from . import case as cs
from . import formats as fm

NT_DEFAULT = 2000


def const_table(k, rho, cp, nT=NT_DEFAULT, nmat=1):
    k = np.broadcast_to(np.asarray(k, np.float64).reshape(-1, 1), (nmat, nT)).copy()
    rho = np.broadcast_to(np.asarray(rho, np.float64).reshape(-1, 1), (nmat, nT)).copy()
    cp = np.broadcast_to(np.asarray(cp, np.float64).reshape(-1, 1), (nmat, nT)).copy()
    h = np.cumsum(rho * cp, axis=1)
    return fm.PropertyTable(1, cp, rho, k, h, [f"mat{m+1}" for m in range(nmat)])


def knot_table(knots_T, knots_v, nT=NT_DEFAULT):
    """Resample knots to 1 K rows T=1..nT by linear interpolation, clamped ends."""
    T = np.arange(1, nT + 1, dtype=np.float64)
    return np.interp(T, knots_T, knots_v)


def nodes_from_axes(x, y, z):
    """Structured node array (nk+1, nj+1, ni+1, 3) from 1-D axis coordinates."""
    x, y, z = (np.asarray(a, np.float64) for a in (x, y, z))
    out = np.empty((z.size, y.size, x.size, 3))
    out[..., 0] = x[None, None, :]
    out[..., 1] = y[None, :, None]
    out[..., 2] = z[:, None, None]
    return out


def stretch(n, L, a, phi):
    """Smooth monotone map of n cells on [0, L]: spacing ~ 1 + a cos(2 pi xi + phi),
    so the largest/smallest spacing ratio is (1+a)/(1-a)."""
    xi = np.linspace(0.0, 1.0, n + 1)
    s = xi + a / (2 * np.pi) * (np.sin(2 * np.pi * xi + phi) - np.sin(phi))
    return L * s


def face_entries(b, dims, f, typ, rrec, ib_extra=None):
    """BC entries of face f of block b (1-based), i/j fastest over the face:
    (ib (n, 15) int32, rb (n, 6) float64).  ib_extra(i, j, k) maps the cell
    index arrays to the (n, 9) donor columns (connections)."""
    ni, nj, nk = dims
    if f <= 2:
        kk, jj = np.meshgrid(np.arange(1, nk + 1), np.arange(1, nj + 1), indexing="ij")
        ii = np.full_like(jj, 1 if f == 1 else ni)
    elif f <= 4:
        kk, ii = np.meshgrid(np.arange(1, nk + 1), np.arange(1, ni + 1), indexing="ij")
        jj = np.full_like(ii, 1 if f == 3 else nj)
    else:
        jj, ii = np.meshgrid(np.arange(1, nj + 1), np.arange(1, ni + 1), indexing="ij")
        kk = np.full_like(ii, 1 if f == 5 else nk)
    ii, jj, kk = ii.reshape(-1), jj.reshape(-1), kk.reshape(-1)
    n = ii.size
    ib = np.zeros((n, fm.BC_NI), np.int32)
    ib[:, 0] = b
    ib[:, 1], ib[:, 2], ib[:, 3] = ii, jj, kk
    ib[:, 4] = f
    ib[:, 5] = typ
    if ib_extra is not None:
        ib[:, 6:15] = ib_extra(ii, jj, kk)
    rb = np.tile(np.asarray(rrec, np.float64), (n, 1))
    return ib, rb


def single_block_bc(b, dims, faces):
    """faces: list of 6 (type, rrec) for faces 1..6."""
    parts = [face_entries(b, dims, f, *faces[f - 1]) for f in range(1, 7)]
    return np.concatenate([p[0] for p in parts]), np.concatenate([p[1] for p in parts])


def bclist(ib, rb):
    ib = np.asarray(ib, np.int32).reshape(-1, fm.BC_NI)
    rb = np.asarray(rb, np.float64).reshape(-1, fm.BC_NR)
    return fm.BCList(ib, rb, [None] * ib.shape[0])


def R(qw=0.0, Tw=0.0, hconv=0.0, qrad=0.0, Tref=0.0, eps=0.0):
    return (qw, Tw, hconv, qrad, Tref, eps)


def plate2d(n=128, L=1.0, k=50.0, rho=7800.0, cp=460.0, T_left=500.0, T_right=300.0, T0=300.0,
            noise=1.0, seed=0, n_rk=3, vnn=0.8, prim=True, irs_beta=0.0, res_threshold=1e-10,
            mgl=1, level_iter=None):
    dx = L / n
    x = np.linspace(0.0, L, n + 1)
    nodes = nodes_from_axes(x, x.copy(), np.array([0.0, dx]))
    rng = np.random.Generator(np.random.PCG64(seed))
    T = (T0 + rng.uniform(-noise, noise, size=n * n)).reshape(1, n, n)
    blk = cs.Block((n, n, 1), nodes, T, np.ones_like(T), np.zeros_like(T))
    faces = [(302, R(Tw=T_left)), (302, R(Tw=T_right)), (301, R(qw=0.0)), (301, R(qw=0.0)),
             (0, R()), (0, R())]
    bcs = []
    for m in range(mgl):
        d = cs.coarse_dims((n, n, 1), mgl)[m]
        bcs.append(bclist(*single_block_bc(1, d, faces)))
    p = cs.Params(n_rk=n_rk, vnn=vnn, prim=prim, irs_beta=irs_beta, mgl=mgl,
                  level_iter=level_iter or [0] * mgl, res_threshold=res_threshold)
    if mgl > 1 and level_iter is None:
        p.level_iter = [10**9] * mgl
    return cs.Case([blk], bcs, const_table(k, rho, cp), p, [], [np.full(b.n, -1, np.int32) for b in bcs],
                   name=f"plate2d_{n}")


def cube_table(seed=1, nT=NT_DEFAULT, rho=7800.0, cp=460.0):
    rng = np.random.Generator(np.random.PCG64(seed))
    kT = np.linspace(250.0, 1500.0, 32)
    kv = (20.0 + 0.02 * (kT - 300.0)) * (1.0 + 0.05 * rng.uniform(-1.0, 1.0, size=32))
    k = knot_table(kT, kv, nT)
    t = const_table(1.0, rho, cp, nT)
    t.kap[0] = k
    return t


def cube3d(n=256, L=0.1, seed=1, qvol=1e6, hconv=50.0, Tinf=300.0, T0=300.0, n_rk=3, vnn=2.0,
           prim=True, irs_beta=0.5, mgl=3, level_iter=None, res_threshold=1e-10, stretch_max=0.6):
    rng = np.random.Generator(np.random.PCG64(seed))
    axes = []
    for _ in range(3):
        a = rng.uniform(0.2, stretch_max)
        phi = rng.uniform(0.0, 2 * np.pi)
        axes.append(stretch(n, L, a, phi))
    nodes = nodes_from_axes(*axes)
    T = np.full((n, n, n), T0)
    blk = cs.Block((n, n, n), nodes, T, np.ones_like(T), np.full_like(T, qvol))
    faces = [(303, R(hconv=hconv, Tref=Tinf))] * 6
    bcs = [bclist(*single_block_bc(1, d, faces)) for d in cs.coarse_dims((n, n, n), mgl)]
    p = cs.Params(n_rk=n_rk, vnn=vnn, prim=prim, irs_beta=irs_beta, mgl=mgl,
                  level_iter=level_iter or [10**9] * mgl, res_threshold=res_threshold)
    return cs.Case([blk], bcs, cube_table(seed), p, [], [np.full(b.n, -1, np.int32) for b in bcs],
                   name=f"cube3d_{n}")


def sweep(ncase=256, lo=32, hi=96, L=0.1, seed=5, n_rk=3, vnn=2.0, prim=True, irs_beta=0.5, mgl=3,
          res_threshold=1e-10, T0=300.0, Tinf=300.0):
    """Independent steady cubes as one multi-block domain: per-case k
    (constant table row per case), h and qvol."""
    rng = np.random.Generator(np.random.PCG64(seed))
    blocks, faces_all = [], []
    ks = rng.uniform(5.0, 200.0, size=ncase)
    hs = rng.uniform(10.0, 500.0, size=ncase)
    qs = rng.uniform(0.0, 1e6, size=ncase)
    for c in range(ncase):
        dims = tuple(int(4 * rng.integers(lo // 4, hi // 4 + 1)) for _ in range(3))
        axes = [np.linspace(0.0, L, d + 1) for d in dims]
        nodes = nodes_from_axes(*axes)
        T = np.full(dims[::-1], T0)
        blocks.append(cs.Block(dims, nodes, T, np.full_like(T, float(c + 1)), np.full_like(T, qs[c])))
        faces_all.append([(303, R(hconv=hs[c], Tref=Tinf))] * 6)
    bcs = []
    for m in range(mgl):
        ib, rb = [], []
        for c, b in enumerate(blocks):
            d = cs.coarse_dims(b.dims, mgl)[m]
            a, r = single_block_bc(c + 1, d, faces_all[c])
            ib.append(a)
            rb.append(r)
        bcs.append(bclist(np.concatenate(ib), np.concatenate(rb)))
    p = cs.Params(n_rk=n_rk, vnn=vnn, prim=prim, irs_beta=irs_beta, mgl=mgl, level_iter=[10**9] * mgl,
                  res_threshold=res_threshold)
    tab = const_table(ks, 7800.0, 460.0, nmat=ncase)
    return cs.Case(blocks, bcs, tab, p, [], [np.full(b.n, -1, np.int32) for b in bcs], name=f"sweep_{ncase}")


def material_tables(nmat=3, seed=3, nT=NT_DEFAULT, npts=16):
    """nmat materials, each with npts-point rho(T), cp(T), k(T) knots on
    [200, 1800] K (seeded), resampled to 1 K rows; h = cumulative rho*cp."""
    rng = np.random.Generator(np.random.PCG64(seed))
    kT = np.linspace(200.0, 1800.0, npts)
    cp, rho, kap = [], [], []
    for _ in range(nmat):
        k0, r0, c0 = rng.uniform(5.0, 150.0), rng.uniform(2000.0, 9000.0), rng.uniform(300.0, 900.0)
        ph = rng.uniform(0.0, 2 * np.pi)
        kap.append(knot_table(kT, k0 * (1.0 + 0.3 * np.sin(2 * np.pi * (kT - 200.0) / 1600.0 + ph)), nT))
        rho.append(knot_table(kT, r0 * (1.0 - 2e-5 * (kT - 300.0)), nT))
        cp.append(knot_table(kT, c0 * (1.0 + 3e-4 * (kT - 300.0)), nT))
    cp, rho, kap = np.array(cp), np.array(rho), np.array(kap)
    return fm.PropertyTable(1, cp, rho, kap, np.cumsum(rho * cp, axis=1), [f"mat{m+1}" for m in range(nmat)])


def flux_plate(n=512, nseg=64, L=1.0, seed=11, nprobe=32, probe_rows=8, n_rk=3, vnn=0.8, T0=300.0,
               q_lo=1e4, q_hi=5e4, mat_seed=3):
    """Transient 2-D plate of the heat-flux inverse problem.

    n x n cells on [0, L]^2; one material with tabulated rho(T), cp(T), k(T)
    (16 knots, resampled to 1 K rows); walls: face 1 302 T = 300 K, face 2
    303 h = 50, face 3 (min j) 301 with the unknown flux q per segment of
    n / nseg cells -- the bc.txt value is -q, which heats through a min face
    (the 301 value is signed along +xi) --, face 4 303 h = 20; T_inf = 300 K.
    Global dt (time-accurate).  Returns (case, info) with info: seg_entries
    (nseg x n/nseg 0-based BC entry indices of face 3), q_true (smooth seeded
    field in [q_lo, q_hi]), probes (nprobe x 4 1-based b,i,j,k, rows j <=
    probe_rows, seeded), kscale_true = 1."""
    rng = np.random.Generator(np.random.PCG64(seed))
    assert n % nseg == 0
    x = np.linspace(0.0, L, n + 1)
    dx = L / n
    nodes = nodes_from_axes(x, x.copy(), np.array([0.0, dx]))
    T = np.full((1, n, n), T0)
    blk = cs.Block((n, n, 1), nodes, T, np.ones_like(T), np.zeros_like(T))
    xs = (np.arange(nseg) + 0.5) / nseg
    f = np.zeros(nseg)
    for kk in range(1, 5):
        f += rng.normal() / kk * np.sin(2 * np.pi * kk * xs + rng.uniform(0, 2 * np.pi))
    f = (f - f.min()) / max(f.max() - f.min(), 1e-30)
    q_true = q_lo + (q_hi - q_lo) * f
    faces = [(302, R(Tw=T0)), (303, R(hconv=50.0, Tref=T0)), (301, R(qw=-q_true.mean())),
             (303, R(hconv=20.0, Tref=T0)), (0, R()), (0, R())]
    ib, rb = single_block_bc(1, (n, n, 1), faces)
    off3 = 2 * n  # faces 1 and 2 hold n entries each
    seg_entries = (off3 + np.arange(n)).reshape(nseg, n // nseg).astype(np.int32)
    for sgi in range(nseg):
        rb[seg_entries[sgi], 0] = -q_true[sgi]
    bcs = [bclist(ib, rb)]
    pi = rng.integers(1, n + 1, size=nprobe)
    pj = rng.integers(1, probe_rows + 1, size=nprobe)
    probes = np.stack([np.ones(nprobe, int), pi, pj, np.ones(nprobe, int)], axis=1).astype(np.int32)
    p = cs.Params(n_rk=n_rk, vnn=vnn, prim=True, irs_beta=0.0, time_accurate=True, mgl=1, level_iter=[10**9])
    tab = material_tables(nmat=1, seed=mat_seed)
    case = cs.Case([blk], bcs, tab, p, [], [np.full(bcs[0].n, -1, np.int32)], name=f"flux_plate_{n}")
    return case, {"seg_entries": seg_entries, "q_true": q_true, "probes": probes, "kscale_true": 1.0}


def lattice(nl=4, lo=16, hi=128, seed=7, h=1e-3, amp=0.05, nmat=3, T0=500.0, n_rk=3, vnn=0.8,
            prim=True, irs_beta=0.0, time_accurate=True, mgl=1, walls=None, res_threshold=1e-10):
    """nl^3 blocks (nl an int) or nx*ny*nz blocks (nl a 3-tuple) in a lattice
    with per-slab dimensions drawn from U{lo..hi} (so neighbouring blocks
    share face dimensions), joined by 1-to-1 101 connections.  Nodes are a smooth global curvilinear map: base spacing h,
    each coordinate perturbed by amp*h*sin(2 pi (k . X/Lg) + phi) (seeded), so
    the faces of neighbouring blocks coincide.  Exterior faces by global side:
    x-min 302, x-max 304, y-min 301, y-max 304, z-min 302, z-max 301 unless
    `walls` overrides them (6 (type, rrec) pairs)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    step = 2 ** (mgl - 1)
    nls = (int(nl),) * 3 if np.isscalar(nl) else tuple(int(x) for x in nl)
    dims = [rng.integers(lo // step, hi // step + 1, size=nls[d]) * step for d in range(3)]
    off = [np.concatenate([[0], np.cumsum(d)]) for d in dims]
    Lg = [off[d][-1] * h for d in range(3)]
    kv = rng.integers(1, 3, size=(3, 3))
    ph = rng.uniform(0.0, 2 * np.pi, size=3)

    def node_map(I, J, K):
        X = np.stack([I * h, J * h, K * h], axis=-1)
        xi = X / np.asarray(Lg)
        out = X.copy()
        for c in range(3):
            out[..., c] += amp * h * np.sin(2 * np.pi * (xi @ kv[c]) + ph[c])
        return out

    walls = walls or [(302, R(Tw=450.0)), (304, R(eps=0.8, Tref=300.0)), (301, R(qw=-5e3)),
                      (304, R(eps=0.8, Tref=300.0)), (302, R(Tw=550.0)), (301, R(qw=0.0))]
    blocks, bid = [], {}
    mats = rng.integers(1, nmat + 1, size=nls[0] * nls[1] * nls[2])
    for c in range(nls[2]):
        for b in range(nls[1]):
            for a in range(nls[0]):
                n = len(blocks)
                bid[(a, b, c)] = n + 1
                ni, nj, nk = int(dims[0][a]), int(dims[1][b]), int(dims[2][c])
                I = off[0][a] + np.arange(ni + 1)
                J = off[1][b] + np.arange(nj + 1)
                K = off[2][c] + np.arange(nk + 1)
                KK, JJ, II = np.meshgrid(K, J, I, indexing="ij")
                nodes = np.ascontiguousarray(node_map(II.astype(float), JJ.astype(float), KK.astype(float)))
                T = np.full((nk, nj, ni), T0)
                blocks.append(cs.Block((ni, nj, nk), nodes, T, np.full_like(T, float(mats[n])), np.zeros_like(T)))

    def lattice_bc(level):
        ib, rb = [], []
        for c in range(nls[2]):
            for b in range(nls[1]):
                for a in range(nls[0]):
                    me = bid[(a, b, c)]
                    d0 = cs.coarse_dims(blocks[me - 1].dims, level + 1)[level]
                    for f in range(1, 7):
                        ax, hi_side = (f - 1) // 2, (f - 1) % 2
                        pos = [a, b, c]
                        pos[ax] += 1 if hi_side else -1
                        if 0 <= pos[ax] < nls[ax]:
                            nb_ = bid[tuple(pos)]
                            dn = cs.coarse_dims(blocks[nb_ - 1].dims, level + 1)[level]
                            fs = f + 1 if f % 2 == 1 else f - 1

                            def extra(i, j, k, ax=ax, hi_side=hi_side, nb_=nb_, dn=dn, fs=fs):
                                q = [i.copy(), j.copy(), k.copy()]
                                q[ax][:] = dn[ax] if not hi_side else 1
                                cols = [np.full_like(i, nb_), q[0], q[1], q[2], np.full_like(i, fs),
                                        np.ones_like(i), np.zeros_like(i), np.zeros_like(i), np.ones_like(i)]
                                return np.stack(cols, axis=1)

                            x, y = face_entries(me, d0, f, 101, R(), extra)
                        else:
                            typ, rr = walls[f - 1]
                            x, y = face_entries(me, d0, f, typ, rr)
                        ib.append(x)
                        rb.append(y)
        return bclist(np.concatenate(ib), np.concatenate(rb))

    bcs = [lattice_bc(m) for m in range(mgl)]
    p = cs.Params(n_rk=n_rk, vnn=vnn, prim=prim, irs_beta=irs_beta, time_accurate=time_accurate, mgl=mgl,
                  level_iter=[10**9] * mgl, res_threshold=res_threshold)
    return cs.Case(blocks, bcs, material_tables(nmat, seed + 1), p, [], [np.full(x.n, -1, np.int32) for x in bcs],
                   name=f"lattice{nl}" if np.isscalar(nl) else f"lattice{'x'.join(map(str, nls))}")


def write_case_dir(case, path, ini_extra=""):
    """Write a Case as a reference case directory (input.ini, INPUT/)."""
    os.makedirs(os.path.join(path, "INPUT"), exist_ok=True)
    zones, stz = [], []
    for b, blk in enumerate(case.blocks):
        ni, nj, nk = blk.dims
        d = {"x": blk.nodes[..., 0], "y": blk.nodes[..., 1], "z": blk.nodes[..., 2], "T": blk.T,
             "matID": blk.matID}
        zones.append((f"B{b+1}-SP", ni + 1, nj + 1, nk + 1, d))
        stz.append((f"B{b+1}-ST", ni + 1, nj + 1, nk + 1, {"x": d["x"], "y": d["y"], "z": d["z"], "qvol": blk.qvol}))
    fm.write_tecplot(os.path.join(path, "INPUT", "ic.tec"), zones, ["x", "y", "z", "T", "matID"], {3, 4})
    fm.write_tecplot(os.path.join(path, "INPUT", "st.tec"), stz, ["x", "y", "z", "qvol"], {3})
    for m, bcl in enumerate(case.bcs):
        fm.write_bc(os.path.join(path, "INPUT", "bc.txt" if m == 0 else f"bc{m+1}.txt"), bcl)
    fm.write_properties(os.path.join(path, "INPUT", "properties.dat"), case.table)
    with open(os.path.join(path, "INPUT", "phase.txt"), "w") as fh:
        fh.write("solid phase\n")
        for m in range(case.table.nmat):
            fh.write(f"mat{m+1} 1\n")
    p = case.params
    scheme = {1: "euler", 2: "RK2", 3: "RK3", 4: "RK4"}[p.n_rk]
    with open(os.path.join(path, "input.ini"), "w") as fh:
        fh.write("[FUSS-Parameters]\nnewrun = true\n")
        fh.write(f"res-threshold = {p.res_threshold!r}\ntime-threshold = {p.time_threshold!r}\n")
        fh.write(f"iter-threshold = {p.iter_threshold}\n\n[FUSS-Numerics]\n")
        fh.write(f"time-scheme = {scheme}\nvnn = {p.vnn!r}\nvnn-rise-threshold = {p.rampa_iter}\n")
        fh.write(f"time-accurate = {'true' if p.time_accurate else 'false'}\n")
        fh.write(f"integration-variables = {'prim' if p.prim else 'cons'}\n")
        fh.write(f"irs = {'true' if p.irs_beta > 0 else 'false'}\nirs-beta = {p.irs_beta!r}\n\n")
        fh.write(f"[FUSS-Multigrid]\nlevels = {p.mgl}\n")
        for m in range(p.mgl):
            fh.write(f"level{m+1}-iter = {p.level_iter[m]}\n")
        fh.write(ini_extra)
