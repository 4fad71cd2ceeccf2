"""Readers and writers for the FUSS on-disk formats.

The formats are the reference's (open-hydra/FUSS) input contract, kept as they
are so that existing case directories run unchanged:

* Tecplot ASCII ``DATAPACKING=BLOCK`` structured zones: ``INPUT/ic.tec`` (nodal
  x, y, z then cell-centred T, matID), ``INPUT/st.tec`` (cell-centred qvol),
  ``OUTPUT/field*.tec`` (src/lib/io/IO_Solution.f90:99-305, read through
  ORION's ``tec_read_structured_multiblock``);
* ``INPUT/bc.txt`` / ``bc<m>.txt`` boundary lists (src/lib/io/IO_BC.f90:121-245):
  one ``b i j k f type`` line per boundary face cell, grouped by (block, face),
  followed by a type-dependent line;
* ``INPUT/properties.dat`` 1 K property tables in Tecplot POINT zones, columns
  T, cp, rho, k, h (src/lib/io/Load_Table.f90:10-75) and ``INPUT/phase.txt``;
* time-series files of time-dependent 301/302 values
  (src/lib/base/Series_Data_m.f90:20-52);
* the INI configuration (src/lib/config/Register_*.f90 for names and defaults).

Everything is returned as numpy arrays in the layouts the C-ABI consumes:
cell arrays i-fastest ``(nk, nj, ni)``, nodes ``(nk+1, nj+1, ni+1, 3)``.
"""
from __future__ import annotations

import os
import re
from dataclasses import dataclass, field

import numpy as np

# BC integer record: b i j k f type bs is js ks fs d11 d12 d21 d22
BC_NI = 15
# BC real record: qw Tw hconv qrad Tref eps
BC_NR = 6

BC_TYPES = (0, 101, 102, 103, 300, 301, 302, 303, 304, 305)


# --------------------------------------------------------------------------
# Tecplot ASCII (BLOCK packing)
# --------------------------------------------------------------------------
@dataclass
class TecZone:
    name: str
    I: int
    J: int
    K: int
    varnames: list
    cell_vars: set
    solutiontime: float
    data: dict = field(default_factory=dict)

    @property
    def cells(self):
        return (self.I - 1, self.J - 1, self.K - 1)


_zone_kv = re.compile(r"(\w+)\s*=\s*(\([^)]*\)|\"[^\"]*\"|[^,]+)")


def _parse_varlocation(text, nvar):
    cell = set()
    m = re.search(r"VARLOCATION\s*=\s*\(([^)]*\))", text)
    if not m:
        return cell
    body = m.group(1)
    for grp, loc in re.findall(r"\[([^\]]*)\]\s*=\s*(\w+)", body):
        if loc.upper().startswith("CELL"):
            for part in grp.split(","):
                if "-" in part:
                    a, b = part.split("-")
                    cell.update(range(int(a) - 1, int(b)))
                else:
                    cell.add(int(part) - 1)
    return cell


def read_tecplot(path):
    """Read a structured multi-zone Tecplot ASCII BLOCK file.

    Returns a list of TecZone; nodal variables are shaped (K, J, I), cell
    variables (K-1, J-1, I-1) (max(1, .) in each axis), both i-fastest.
    """
    with open(path, "r") as fh:
        text = fh.read()
    lines = text.splitlines()
    varnames = None
    zones = []
    chunks = []
    cur = None
    for ln in lines:
        s = ln.strip()
        if not s:
            continue
        up = s.upper()
        if up.startswith("VARIABLES"):
            varnames = re.findall(r"\"([^\"]*)\"|(\S+)", s.split("=", 1)[1])
            varnames = [a or b for a, b in varnames]
            varnames = [v.strip(",") for v in varnames if v.strip(",")]
            continue
        if up.startswith("TITLE"):
            continue
        if up.startswith("ZONE"):
            if cur is not None:
                zones.append((cur, chunks))
            kv = {}
            for k, v in _zone_kv.findall(s[4:]):
                kv[k.upper()] = v.strip()
            cur = TecZone(
                name=kv.get("T", "").strip('"').strip(),
                I=int(kv.get("I", 1)),
                J=int(kv.get("J", 1)),
                K=int(kv.get("K", 1)),
                varnames=list(varnames or []),
                cell_vars=_parse_varlocation(s, len(varnames or [])),
                solutiontime=float(kv.get("SOLUTIONTIME", "0").replace("D", "E")),
            )
            chunks = []
            continue
        chunks.append(ln)
    if cur is not None:
        zones.append((cur, chunks))
    out = []
    for z, ch in zones:
        vals = np.array(" ".join(ch).replace(",", " ").split(), dtype=np.float64)
        nn = z.I * z.J * z.K
        nc = max(1, z.I - 1) * max(1, z.J - 1) * max(1, z.K - 1)
        pos = 0
        for v, name in enumerate(z.varnames):
            if v in z.cell_vars:
                arr = vals[pos:pos + nc].reshape(max(1, z.K - 1), max(1, z.J - 1), max(1, z.I - 1))
                pos += nc
            else:
                arr = vals[pos:pos + nn].reshape(z.K, z.J, z.I)
                pos += nn
            z.data[name] = arr
        if pos != vals.size:
            raise ValueError(f"{path}: zone {z.name}: expected {pos} values, found {vals.size}")
        out.append(z)
    return out


def _fmt15(a):
    return "\n".join(f" {x: .16E}" for x in np.asarray(a, dtype=np.float64).ravel())


def write_tecplot(path, zones, varnames, cell_vars, solutiontime=0.0):
    """Write zones (list of (name, I, J, K, {var: array})) as Tecplot ASCII BLOCK."""
    nv = len(varnames)
    loc = []
    nodal = [i + 1 for i in range(nv) if i not in cell_vars]
    cells = [i + 1 for i in range(nv) if i in cell_vars]

    def rng(ix):
        return f"{ix[0]}-{ix[-1]}" if len(ix) > 1 else f"{ix[0]}"

    if nodal:
        loc.append(f"[{rng(nodal)}]=NODAL")
    if cells:
        loc.append(f"[{rng(cells)}]=CELLCENTERED")
    with open(path, "w") as fh:
        fh.write(" VARIABLES =" + " ".join(f'"{v}"' for v in varnames) + "\n")
        for name, I, J, K, data in zones:
            fh.write(f" ZONE  T = {name}, I={I}, J={J}, K={K}, DATAPACKING=BLOCK, "
                     f"VARLOCATION=({','.join(loc)}), SOLUTIONTIME={solutiontime: .15E}, STRANDID = 0\n")
            for v in varnames:
                fh.write(_fmt15(data[v]) + "\n")


WALL_TYPES = (301, 302, 303, 304, 305)


def wall_zones(bcl, dims, Tw, qw):
    """Per-entry wall values (fuss_wall / Extract_Properties_From_BC order) to
    the reference's wall zones (Initialize_Wall_File, IO_Wall.f90:12-95): one
    zone per (block, face) that holds a wall entry (its viscous_flag), named
    B<b>F<f>, Tw and qw cell-centred on the face, shaped (K, J, I) with the
    face's own axis of length 1.  dims[b-1] = (ni, nj, nk).

    Returns {(b, f): {"Tw": array, "qw": array}} in block, face order."""
    out = {}
    for n in range(bcl.n):
        b, i, j, k, f, typ = (int(x) for x in bcl.ib[n, :6])
        if typ not in WALL_TYPES:
            continue
        if (b, f) not in out:
            ni, nj, nk = dims[b - 1]
            shape = (nk, nj, 1) if f <= 2 else ((nk, 1, ni) if f <= 4 else (1, nj, ni))
            out[(b, f)] = {"Tw": np.zeros(shape), "qw": np.zeros(shape)}
        z = out[(b, f)]
        at = (k - 1, j - 1, 0) if f <= 2 else ((k - 1, 0, i - 1) if f <= 4 else (0, j - 1, i - 1))
        z["Tw"][at] = Tw[n]
        z["qw"][at] = qw[n]
    return dict(sorted(out.items()))


def write_wall_tecplot(path, nodes, dims, zones, solutiontime=0.0):
    """Write wall zones (wall_zones) as the reference's wall file
    (Write_Wall_Solution, IO_Wall.f90:98-126): x, y, z of the face's nodes,
    then Tw, qw cell-centred.  nodes[b-1]: the block's (nk+1, nj+1, ni+1, 3)
    node array."""
    out = []
    for (b, f), z in zones.items():
        ni, nj, nk = dims[b - 1]
        X = np.asarray(nodes[b - 1]).reshape(nk + 1, nj + 1, ni + 1, 3)
        if f <= 2:
            sl = X[:, :, 0:1] if f == 1 else X[:, :, ni:ni + 1]
        elif f <= 4:
            sl = X[:, 0:1, :] if f == 3 else X[:, nj:nj + 1, :]
        else:
            sl = X[0:1] if f == 5 else X[nk:nk + 1]
        K, J, I = sl.shape[:3]
        out.append((f"B{b}F{f}", I, J, K, {"x": sl[..., 0], "y": sl[..., 1], "z": sl[..., 2],
                                           "Tw": z["Tw"], "qw": z["qw"]}))
    write_tecplot(path, out, ["x", "y", "z", "Tw", "qw"], {3, 4}, solutiontime)


# --------------------------------------------------------------------------
# boundary conditions
# --------------------------------------------------------------------------
@dataclass
class BCList:
    ib: np.ndarray          # (n, 15) int32
    rb: np.ndarray          # (n, 6) float64
    series_file: list       # per entry: None or the time-series file name
    # chimera (102) entries: BC index -> (n1, donors (n1+n2, 4) int32 as
    # (b, i, j, k), volume fractions (n1+n2,)); rows 1..n1 make the first
    # ghost layer, the rest the second (IO_BC.f90:218-229)
    chim: dict = field(default_factory=dict)

    @property
    def n(self):
        return self.ib.shape[0]

    def chimera_arrays(self):
        """Flattened chimera donor lists in BC order: (entry, n1, n2, donors
        (m, 4), vf (m,)) -- the layout of fuss_set_chimera."""
        ent = sorted(self.chim)
        n1 = np.array([self.chim[e][0] for e in ent], np.int32)
        n2 = np.array([len(self.chim[e][2]) - self.chim[e][0] for e in ent], np.int32)
        don = (np.concatenate([self.chim[e][1] for e in ent]).astype(np.int32) if ent
               else np.zeros((0, 4), np.int32))
        vf = np.concatenate([self.chim[e][2] for e in ent]) if ent else np.zeros(0)
        return np.array(ent, np.int32), n1, n2, np.ascontiguousarray(don), np.ascontiguousarray(vf, np.float64)


def _tokens(line):
    return [t for t in re.split(r"[,\s]+", line.strip()) if t]


def _fnum(tok):
    return float(tok.replace("D", "E").replace("d", "e"))


def read_bc(path):
    """Read a bc.txt file (src/lib/io/IO_BC.f90:121-245)."""
    with open(path, "r") as fh:
        lines = [ln for ln in fh.read().splitlines() if ln.strip()]
    ib, rb, sf = [], [], []
    chim = {}
    p = 0
    while p < len(lines):
        t = _tokens(lines[p])
        p += 1
        if len(t) < 6:
            raise ValueError(f"{path}: malformed BC line {p}: {lines[p-1]!r}")
        b, i, j, k, f, typ = (int(x) for x in t[:6])
        irec = [b, i, j, k, f, typ] + [0] * 9
        rrec = [0.0] * 6
        series = None
        if typ in (101, 103):
            u = [int(x) for x in _tokens(lines[p])[:9]]
            p += 1
            irec[6:15] = u
        elif typ in (301, 302):
            tok = _tokens(lines[p])[0]
            p += 1
            try:
                val = _fnum(tok)
            except ValueError:
                val = 0.0
                series = tok
            rrec[0 if typ == 301 else 1] = val
        elif typ == 303:
            u = [_fnum(x) for x in _tokens(lines[p])[:3]]
            p += 1
            rrec[2], rrec[3], rrec[4] = u
        elif typ == 304:
            u = [_fnum(x) for x in _tokens(lines[p])[:2]]
            p += 1
            rrec[5], rrec[4] = u
        elif typ == 305:
            u = [_fnum(x) for x in _tokens(lines[p])[:3]]
            p += 1
            rrec[2], rrec[5], rrec[4] = u
        elif typ == 102:
            # receiver counts of the two ghost layers, then one line per donor:
            # b i j k volume_fraction
            n1, n2 = (int(x) for x in _tokens(lines[p])[:2])
            p += 1
            don, vf = [], []
            for _ in range(n1 + n2):
                u = _tokens(lines[p])
                p += 1
                don.append([int(x) for x in u[:4]])
                vf.append(_fnum(u[4]))
            chim[len(ib)] = (n1, np.asarray(don, np.int32).reshape(-1, 4), np.asarray(vf, np.float64))
        elif typ not in (0, 300):
            raise ValueError(f"{path}: unknown BC type {typ}")
        ib.append(irec)
        rb.append(rrec)
        sf.append(series)
    return BCList(np.asarray(ib, dtype=np.int32).reshape(-1, BC_NI),
                  np.asarray(rb, dtype=np.float64).reshape(-1, BC_NR), sf, chim)


def write_bc(path, bcl: BCList):
    """Write a bc.txt file in the reference layout (I8 integers, E13.6 reals)."""
    out = []
    for n in range(bcl.n):
        q = bcl.ib[n]
        r = bcl.rb[n]
        typ = int(q[5])
        out.append("".join(f"{int(x):8d}" for x in q[:6]))
        if typ in (101, 103):
            out.append("".join(f"{int(x):8d}" for x in q[6:15]))
        elif typ == 301:
            out.append(f" {bcl.series_file[n]}," if bcl.series_file[n] else f"  {r[0]: .15E},")
        elif typ == 302:
            out.append(f" {bcl.series_file[n]}," if bcl.series_file[n] else f"  {r[1]: .15E},")
        elif typ == 303:
            out.append(f"  {r[2]: .15E},  {r[3]: .15E},  {r[4]: .15E},")
        elif typ == 304:
            out.append(f"  {r[5]: .15E},  {r[4]: .15E},")
        elif typ == 305:
            out.append(f"  {r[2]: .15E},  {r[5]: .15E},  {r[4]: .15E},")
        elif typ == 102:
            n1, don, vf = bcl.chim[n]
            out.append(f"{n1:8d}{len(vf) - n1:8d}")
            for d, v in zip(don, vf):
                out.append("".join(f"{int(x):8d}" for x in d) + f"  {v: .15E}")
    with open(path, "w") as fh:
        fh.write("\n".join(out) + "\n")


def check_bc_count(bcl: BCList, dims):
    """Entry-count check of Check_BC (src/lib/io/IO_BC.f90:69-118, :54-62)."""
    n = 0
    for ni, nj, nk in dims:
        n += 2 * nj * nk + 2 * ni * nk + 2 * nj * ni
    if n != bcl.n:
        raise ValueError(f"Boundary conditions number ({bcl.n}) is different than the one "
                         f"of the initial conditions ({n})")


# --------------------------------------------------------------------------
# property tables
# --------------------------------------------------------------------------
@dataclass
class PropertyTable:
    Ti1: int
    cp: np.ndarray    # (nmat, nT)
    rho: np.ndarray
    kap: np.ndarray
    h: np.ndarray
    names: list

    @property
    def nmat(self):
        return self.cp.shape[0]

    @property
    def nT(self):
        return self.cp.shape[1]


def read_properties(path, phase_path=None):
    """Read properties.dat (Tecplot POINT zones of T, cp, rho, k, h)."""
    zones = []
    names = []
    cur = []
    with open(path, "r") as fh:
        for ln in fh:
            s = ln.strip()
            if not s:
                continue
            u = s.upper()
            if u.startswith("TITLE") or u.startswith("VARIABLES"):
                continue
            if u.startswith("ZONE"):
                if cur:
                    zones.append(cur)
                cur = []
                m = re.search(r"T\s*=\s*\"([^\"]*)\"", s)
                names.append(m.group(1) if m else "")
                continue
            if re.match(r"^I\s*=", s, re.I):
                continue
            cur.append(s)
    if cur:
        zones.append(cur)
    tabs = [np.array(" ".join(z).split(), dtype=np.float64).reshape(-1, 5) for z in zones]
    nT = tabs[0].shape[0]
    if any(t.shape[0] != nT for t in tabs):
        raise ValueError("property zones differ in length")
    Ti1 = int(round(tabs[0][0, 0]))
    if not np.allclose(tabs[0][:, 0], Ti1 + np.arange(nT)):
        raise ValueError("property table rows must be 1 K apart")
    if phase_path is not None:
        names = read_phase(phase_path) or names
    return PropertyTable(Ti1, np.stack([t[:, 1] for t in tabs]), np.stack([t[:, 2] for t in tabs]),
                         np.stack([t[:, 3] for t in tabs]), np.stack([t[:, 4] for t in tabs]), names)


def write_properties(path, tab: PropertyTable):
    with open(path, "w") as fh:
        fh.write('TITLE = "Mass Thermodynamic Properties"\n')
        fh.write('VARIABLES = "Temperature", "Cp", "Density", "Conductivity", "Energy"\n')
        for m in range(tab.nmat):
            nm = tab.names[m] if m < len(tab.names) else f"mat{m+1}"
            fh.write(f'ZONE T="{nm}"\nI={tab.nT}, F=POINT\n')
            T = tab.Ti1 + np.arange(tab.nT)
            for q in range(tab.nT):
                fh.write(f"{T[q]:.1f} {float(tab.cp[m, q])!r} {float(tab.rho[m, q])!r} {float(tab.kap[m, q])!r} {float(tab.h[m, q])!r}\n")


def read_phase(path):
    with open(path) as fh:
        rows = [ln.split() for ln in fh.read().splitlines()[1:] if ln.strip()]
    return [r[0] for r in rows]


def enthalpy_column(Ti1, cp, rho):
    """h(T) = cumulative sum of rho*cp over the 1 K rows, as the reference's
    table generator writes it (e.g. test/steady-state/*/INPUT/properties.dat:
    h(1) = rho*cp, h(n) = n*rho*cp for constant properties)."""
    return np.cumsum(rho * cp, axis=-1) + (Ti1 - 1) * 0.0


# --------------------------------------------------------------------------
# time series
# --------------------------------------------------------------------------
def read_series(path):
    """Read a time-series file (Series_Data_m.f90:20-52): optional 'periodic'
    first line, then 'time value' rows."""
    periodic = False
    rows = []
    with open(path) as fh:
        for ln in fh:
            if "periodic" in ln:
                periodic = True
                continue
            t = _tokens(ln)
            if len(t) >= 2:
                rows.append((_fnum(t[0]), _fnum(t[1])))
    a = np.asarray(rows, dtype=np.float64).reshape(-1, 2)
    return periodic, a[:, 0].copy(), a[:, 1].copy()


# --------------------------------------------------------------------------
# INI
# --------------------------------------------------------------------------
def read_ini(path):
    """Minimal INI reader for the FUSS sections: returns {section: {key: str}}."""
    out = {}
    sec = None
    with open(path) as fh:
        for ln in fh:
            s = ln.strip()
            if not s or s[0] in ";#!":
                continue
            m = re.match(r"^\[(.+)\]$", s)
            if m:
                sec = m.group(1).strip()
                out.setdefault(sec, {})
                continue
            if "=" in s and sec is not None:
                k, v = s.split("=", 1)
                out[sec][k.strip()] = v.split(";")[0].strip()
    return out


def ini_get(ini, section, key, default, kind=str):
    v = ini.get(section, {}).get(key, None)
    if v is None:
        v = default
    if kind is bool:
        return str(v).strip().lower().strip(".") in ("true", "t", "1", "yes")
    if kind is float:
        return float(str(v).replace("d", "e").replace("D", "E"))
    if kind is int:
        return int(float(v))
    return str(v)


def exists(path):
    return os.path.isfile(path)
