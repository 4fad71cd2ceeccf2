"""Convert the reference's test data into compact fixtures (run here, where
/root/reference exists; the fixtures travel with the repository).

Each fixture holds the inputs of one reference case exactly as its files give
them (mesh + IC from INPUT/ic.tec, qvol from INPUT/st.tec, bc*.txt, the 1 K
property tables, time series, INI numerics) and the reference's stored
outputs for it:

  temperature_Plate, qvol_Plate, convective_Plate, radiative_Rod
      test/steady-state/<case>/SOLUTION/field.tec -> final T, stop iteration
      test/steady-state/<case>/SOLUTION/wall.tec  -> wall Tw, qw per face zone
  multimat_Plate
      test/transient/multimat_Plate/SOLUTION/field{1,21,41,61,81,101}.tec
  oscillating_Rod
      test/transient/oscillating_Rod/SOLUTION/p1.txt, p2.txt (probe rows)
  mg_irs
      test/numerics-features/residual_convergence/SOLUTION/mg_irs.dat
      (residual of all 80603 iterations, 10 significant digits)
  chimera
      test/numerics-features/chimera/SOLUTION/field.tec -> final T, stop
      iteration; SOLUTION/wall.tec -> wall Tw, qw (two 2-D blocks of
      different j resolution joined by chimera (102) entries)

Usage: python tests/golden/make_fixtures.py [/root/reference]
"""
from __future__ import annotations

import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", ".."))

from fuss_amd import case as cs  # noqa: E402
from fuss_amd import formats as fm  # noqa: E402

CASES = {
    "temperature_Plate": "test/steady-state/temperature_Plate",
    "qvol_Plate": "test/steady-state/qvol_Plate",
    "convective_Plate": "test/steady-state/convective_Plate",
    "radiative_Rod": "test/steady-state/radiative_Rod",
    "multimat_Plate": "test/transient/multimat_Plate",
    "oscillating_Rod": "test/transient/oscillating_Rod",
    "mg_irs": "test/numerics-features/residual_convergence/mg_irs",
    "chimera": "test/numerics-features/chimera",
}


def case_arrays(c):
    out = {}
    out["nb"] = np.int32(c.nb)
    for b, blk in enumerate(c.blocks):
        out[f"b{b}_dims"] = np.asarray(blk.dims, np.int32)
        out[f"b{b}_nodes"] = blk.nodes
        out[f"b{b}_T"] = blk.T
        out[f"b{b}_matID"] = blk.matID
        out[f"b{b}_qvol"] = blk.qvol
    for m, bcl in enumerate(c.bcs):
        out[f"bc{m}_ib"] = bcl.ib
        out[f"bc{m}_rb"] = bcl.rb
        out[f"bc{m}_sid"] = c.series_id[m]
        if bcl.chim:
            ent, n1, n2, don, vf = bcl.chimera_arrays()
            out[f"bc{m}_chent"], out[f"bc{m}_chn1"], out[f"bc{m}_chn2"] = ent, n1, n2
            out[f"bc{m}_chdon"], out[f"bc{m}_chvf"] = don, vf
    out["nseries"] = np.int32(len(c.series))
    for s, (per, t, v) in enumerate(c.series):
        out[f"s{s}_periodic"] = np.int32(per)
        out[f"s{s}_t"] = t
        out[f"s{s}_v"] = v
    t = c.table
    out["tab_Ti1"] = np.int32(t.Ti1)
    out["tab_cp"], out["tab_rho"], out["tab_kap"], out["tab_h"] = t.cp, t.rho, t.kap, t.h
    out["params"] = np.frombuffer(json.dumps(c.params.__dict__).encode(), dtype=np.uint8)
    return out


def main(ref="/root/reference"):
    for name, rel in CASES.items():
        d = os.path.join(ref, rel)
        c = cs.load_case_dir(d)
        arr = case_arrays(c)
        if name in ("temperature_Plate", "qvol_Plate", "convective_Plate", "radiative_Rod", "chimera"):
            sol = fm.read_tecplot(os.path.join(d, "SOLUTION", "field.tec"))
            arr["stop_iter"] = np.int32(round(sol[0].solutiontime))
            for b, z in enumerate(sol):
                arr[f"ref_T{b}"] = z.data["T"]
            # wall output of the same state (Tw, qw per wall face zone)
            for z in fm.read_tecplot(os.path.join(d, "SOLUTION", "wall.tec")):
                arr[f"wall_{z.name}_Tw"] = z.data["Tw"]
                arr[f"wall_{z.name}_qw"] = z.data["qw"]
        elif name == "multimat_Plate":
            snaps = [1, 21, 41, 61, 81, 101]
            arr["snap_iter"] = np.asarray([10 * s for s in snaps], np.int32)
            for q, s in enumerate(snaps):
                sol = fm.read_tecplot(os.path.join(d, "SOLUTION", f"field{s}.tec"))
                arr[f"snap{q}_time"] = np.float64(sol[0].solutiontime)
                for b, z in enumerate(sol):
                    arr[f"snap{q}_T{b}"] = z.data["T"]
        elif name == "oscillating_Rod":
            arr["p1"] = np.loadtxt(os.path.join(d, "SOLUTION", "p1.txt"))
            arr["p2"] = np.loadtxt(os.path.join(d, "SOLUTION", "p2.txt"))
            arr["probes"] = np.asarray([[1, 90, 1, 1], [1, 91, 1, 1]], np.int32)
        elif name == "mg_irs":
            h = np.loadtxt(os.path.join(ref, "test/numerics-features/residual_convergence/SOLUTION/mg_irs.dat"))
            assert np.array_equal(h[:, 0], np.arange(1, h.shape[0] + 1))
            assert np.all(h[:, 1] == -1.0)
            arr["res_hist"] = h[:, 2].copy()
        out = os.path.join(HERE, f"{name}.npz")
        np.savez_compressed(out, **arr)
        print(f"{out}: {os.path.getsize(out)/1024:.0f} KiB")


if __name__ == "__main__":
    main(*sys.argv[1:])
