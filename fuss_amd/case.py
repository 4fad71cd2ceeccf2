"""Case description: the flattened inputs of one FUSS run.

A ``Case`` carries exactly what the reference's FUSS_setup produces before the
first Explicit_Step (src/lib/driver/Wrap_Setup.f90:10-165): block dimensions
and nodes, the initial T / matID / qvol fields, the BC list of every multigrid
level, the 1 K property tables, time series of time-dependent wall values and
the numerics parameters of the INI file.  ``load_case_dir`` reads a reference
case directory (input.ini, INPUT/, MESH/) unchanged.
"""
from __future__ import annotations

import os
from dataclasses import dataclass, field

import numpy as np

from . import formats as fm


@dataclass
class Block:
    dims: tuple           # (ni, nj, nk) interior cells
    nodes: np.ndarray     # (nk+1, nj+1, ni+1, 3)
    T: np.ndarray         # (nk, nj, ni)
    matID: np.ndarray     # (nk, nj, ni) float64 material number (1-based)
    qvol: np.ndarray      # (nk, nj, ni)

    @property
    def ncell(self):
        ni, nj, nk = self.dims
        return ni * nj * nk


@dataclass
class Params:
    """Numerics of the INI file (src/lib/config/Register_Numerics.f90:30-48,
    Register_Sim_Param.f90:121-124, Register_IO.f90:31)."""
    n_rk: int = 1                 # euler=1, RK2=2, RK3=3, RK4=4 (Jameson four-stage, no reference counterpart)
    vnn: float = 0.3
    rampa_iter: int = 0           # vnn-rise-threshold
    time_accurate: bool = False
    prim: bool = False            # integration-variables: cons (default) / prim
    irs_beta: float = 0.0         # IRS enabled when > 0
    irs_mode: int = 0             # 0 jacobi2 (reference), 1 tridiag (exact line solve, no reference counterpart)
    mgl: int = 1                  # multigrid levels
    level_iter: list = field(default_factory=lambda: [0])
    res_threshold: float = 1e-10
    time_threshold: float = 1e30
    iter_threshold: int = 1000000000
    res_diter: int = 1

    def itermax(self):
        """Per-level iteration limits as FUSS_setup leaves them: level 1 gets
        iter-threshold unless multigrid overrides it with level1-iter
        (Wrap_Setup.f90:49, Mod_Multigrid.f90:346-348)."""
        if self.mgl == 1:
            return [self.iter_threshold]
        return [int(self.level_iter[m]) for m in range(self.mgl)]


@dataclass
class Case:
    blocks: list
    bcs: list                     # one formats.BCList per level
    table: fm.PropertyTable
    params: Params
    series: list = field(default_factory=list)   # (periodic, t, v)
    series_id: list = field(default_factory=list)  # per level: int array per entry (-1 none)
    name: str = ""

    @property
    def nb(self):
        return len(self.blocks)

    @property
    def ndir(self):
        """Check_Mesh_Type (src/lib/numerics/space/Lib_Metrics.f90:13-66): from block 1."""
        ni, nj, nk = self.blocks[0].dims
        if nk > 1:
            return 3
        if nj == 1:
            return 1
        return 2

    @property
    def ncell(self):
        return sum(b.ncell for b in self.blocks)


def params_from_ini(ini):
    g = fm.ini_get
    p = Params()
    scheme = g(ini, "FUSS-Numerics", "time-scheme", "euler")
    p.n_rk = 1 if scheme == "euler" else int(scheme[2])
    p.vnn = g(ini, "FUSS-Numerics", "vnn", "0.3", float)
    p.rampa_iter = g(ini, "FUSS-Numerics", "vnn-rise-threshold", "0", int)
    p.time_accurate = g(ini, "FUSS-Numerics", "time-accurate", ".false.", bool)
    p.prim = g(ini, "FUSS-Numerics", "integration-variables", "cons") == "prim"
    irs = g(ini, "FUSS-Numerics", "irs", ".false.", bool)
    beta = g(ini, "FUSS-Numerics", "irs-beta", "0.0", float)
    p.irs_beta = beta if (irs or beta > 0.0) else 0.0
    # extension key (absent from the reference's registry): jacobi2 | tridiag
    p.irs_mode = 1 if g(ini, "FUSS-Numerics", "irs-mode", "jacobi2") == "tridiag" else 0
    p.mgl = max(1, g(ini, "FUSS-Multigrid", "levels", "1", int))
    p.level_iter = [g(ini, "FUSS-Multigrid", f"level{m+1}-iter", "0", int) for m in range(p.mgl)]
    p.res_threshold = g(ini, "FUSS-Parameters", "res-threshold", "1e-10", float)
    p.time_threshold = g(ini, "FUSS-Parameters", "time-threshold", "1e30", float)
    p.iter_threshold = g(ini, "FUSS-Parameters", "iter-threshold", "1000000000", int)
    p.res_diter = g(ini, "FUSS-IO", "res-diter", "1", int)
    return p


def load_case_dir(path, ini_name="input.ini"):
    """Read a reference case directory without modification."""
    ini = fm.read_ini(os.path.join(path, ini_name))
    params = params_from_ini(ini)
    ic = fm.read_tecplot(os.path.join(path, "INPUT", "ic.tec"))
    st_path = os.path.join(path, "INPUT", "st.tec")
    st = fm.read_tecplot(st_path) if os.path.isfile(st_path) else None
    blocks = []
    for zi, z in enumerate(ic):
        ni, nj, nk = max(1, z.I - 1), max(1, z.J - 1), max(1, z.K - 1)
        nodes = np.stack([z.data[v] for v in z.varnames[:3]], axis=-1)
        T = z.data[z.varnames[3]].astype(np.float64)
        mat = z.data[z.varnames[4]].astype(np.float64)
        if st is not None:
            qz = st[zi]
            qvol = qz.data[qz.varnames[-1]].astype(np.float64)
        else:
            qvol = np.zeros_like(T)
        blocks.append(Block((ni, nj, nk), np.ascontiguousarray(nodes), np.ascontiguousarray(T),
                            np.ascontiguousarray(mat), np.ascontiguousarray(qvol)))
    table = fm.read_properties(os.path.join(path, "INPUT", "properties.dat"),
                               os.path.join(path, "INPUT", "phase.txt"))
    bcs = []
    series, sid_all, sidx = [], [], {}
    for m in range(params.mgl):
        name = "bc.txt" if m == 0 else f"bc{m+1}.txt"
        bcl = fm.read_bc(os.path.join(path, "INPUT", name))
        sid = np.full(bcl.n, -1, dtype=np.int32)
        for n, s in enumerate(bcl.series_file):
            if s is None:
                continue
            if s not in sidx:
                sidx[s] = len(series)
                series.append(fm.read_series(os.path.join(path, s)))
            sid[n] = sidx[s]
        bcs.append(bcl)
        sid_all.append(sid)
    fm.check_bc_count(bcs[0], [b.dims for b in blocks])
    return Case(blocks, bcs, table, params, series, sid_all, name=os.path.basename(path.rstrip("/")))


def coarse_dims(dims, levels):
    """Block dimensions of every multigrid level (Mod_Multigrid.f90:324-328)."""
    out = [tuple(dims)]
    for _ in range(1, levels):
        ni, nj, nk = out[-1]
        out.append((ni // 2, nj // 2, max(1, nk // 2)))
    return out
