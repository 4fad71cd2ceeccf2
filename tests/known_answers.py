"""Published known answers of the reference's verification cases, evaluated
on a converged (or transient) state the way the reference's verification
scripts do, with their pass bounds:

* NAFEMS steady benchmarks (targets and bounds as stated in
  test/steady-state/*/verify.py): the temperature at a point, bilinear
  between the four surrounding cell centres (temperature_Plate T(0.2, 0.2) =
  260.5 C, qvol_Plate T(0.3, 0.2) = 310.1 C), the wall temperature of face 2
  of block 1 near a height (convective_Plate: mean of the two face cells
  nearest y = 0.2, 18.25 C) or over the face (radiative_Rod: 927 K); < 1 %.
* multimat_Plate (test/transient/multimat_Plate/verify.py:43-58, 243): the
  normalised temperature of block 2's central column against the Fourier
  series of a slab between 101 K (y = 0) and 100 K (y = 2 m), diffusivity 1,
  100 terms; max |error| < 0.02 at every snapshot.
* oscillating_Rod (test/transient/oscillating_Rod/verify.py:18-19, 115-146):
  the peak of the mean of the two probes above 2000 K, 94.69 K; < 1 %.
"""
import numpy as np

C0 = 273.15

STEADY_TARGETS = {
    "temperature_Plate": ("point", (0.2, 0.2), C0 + 260.5),
    "qvol_Plate": ("point", (0.3, 0.2), C0 + 310.1),
    "convective_Plate": ("wall_near", 0.2, C0 + 18.25),
    "radiative_Rod": ("wall_mean", None, 927.0),
}


def _centres(nodes):
    """Cell centres (nj, ni, 2) of the k = 0 node layer of a 2-D block."""
    n = nodes[0]
    return 0.25 * (n[:-1, :-1, :2] + n[:-1, 1:, :2] + n[1:, :-1, :2] + n[1:, 1:, :2])


def point_T(nodes, T, x, y):
    """Bilinear interpolation of cell-centred T (k = 0 plane) at (x, y) between
    the four surrounding centres of a Cartesian block."""
    cc = _centres(nodes)
    xu, yu = np.unique(cc[0, :, 0]), np.unique(cc[:, 0, 1])
    ix = int(np.clip(np.searchsorted(xu, x) - 1, 0, xu.size - 2))
    iy = int(np.clip(np.searchsorted(yu, y) - 1, 0, yu.size - 2))
    tx = (x - xu[ix]) / (xu[ix + 1] - xu[ix])
    ty = (y - yu[iy]) / (yu[iy + 1] - yu[iy])
    T2 = T[0]
    return ((1 - tx) * (1 - ty) * T2[iy, ix] + tx * (1 - ty) * T2[iy, ix + 1]
            + (1 - tx) * ty * T2[iy + 1, ix] + tx * ty * T2[iy + 1, ix + 1])


def steady_value(case, name, T_blocks, wall_Tw):
    """The case's known-answer quantity from the block fields and the wall
    output (per BC entry), or None when the case has no target."""
    if name not in STEADY_TARGETS:
        return None
    kind, where, _ = STEADY_TARGETS[name]
    b0 = case.blocks[0]
    if kind == "point":
        return point_T(b0.nodes, T_blocks[0], *where)
    ib = case.bcs[0].ib
    sel = (ib[:, 0] == 1) & (ib[:, 4] == 2) & (ib[:, 3] == 1)
    Tw = np.asarray(wall_Tw)[sel]
    if kind == "wall_mean":
        return float(np.mean(Tw))
    # wall cells of face 2 (i = ni) by j: centres between consecutive face nodes
    j = ib[sel, 2]
    ynode = b0.nodes[0, :, -1, 1]
    yc = 0.5 * (ynode[j - 1] + ynode[j])
    near = np.argsort(np.abs(yc - where))[:2]
    return float(np.mean(Tw[near]))


def check_steady(case, name, T_blocks, wall_Tw):
    v = steady_value(case, name, T_blocks, wall_Tw)
    if v is None:
        return None
    ref = STEADY_TARGETS[name][2]
    err = abs(v - ref) / ref
    assert err < 0.01, (name, v, ref, err)
    return err


def slab_series(y, t, L=2.0, alpha=1.0, T0=1.0, TL=0.0, terms=100):
    """Normalised transient slab temperature: steady line plus the Fourier
    series of the initial step (T = 0 inside, T0 at y = 0, TL at y = L)."""
    y = np.asarray(y, float)
    T = T0 + y / L * (TL - T0)
    for r in range(1, terms + 1):
        A = 2.0 / (r * np.pi) * ((-1) ** r * TL - T0)
        T = T + A * np.sin(r * np.pi * y / L) * np.exp(-r * r * np.pi ** 2 * alpha * t / L ** 2)
    return T


def multimat_error(case, T_block2, t):
    """Max |normalised T - series| along block 2's central i column."""
    b = case.blocks[1]
    ni = b.dims[0]
    ic = ni // 2
    ynode = b.nodes[0, :, ic, 1]
    yc = 0.5 * (ynode[:-1] + ynode[1:])
    Tn = (T_block2[0, :, ic] - 100.0) / (101.0 - 100.0)
    return float(np.max(np.abs(Tn - slab_series(yc, t))))


OSC_BASE, OSC_DT = 2000.0, 94.69


def oscillating_error(p1, p2):
    """Relative error of the peak mean-probe rise above 2000 K."""
    dT = float(np.max(0.5 * (np.asarray(p1) + np.asarray(p2)))) - OSC_BASE
    return abs(dT - OSC_DT) / OSC_DT
