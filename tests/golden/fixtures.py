"""Load the golden fixtures written by make_fixtures.py back into Cases."""
from __future__ import annotations

import json
import os

import numpy as np

from fuss_amd import case as cs
from fuss_amd import formats as fm

HERE = os.path.dirname(os.path.abspath(__file__))
NAMES = ("temperature_Plate", "qvol_Plate", "convective_Plate", "radiative_Rod", "multimat_Plate",
         "oscillating_Rod", "mg_irs", "chimera")


def load(name):
    z = np.load(os.path.join(HERE, f"{name}.npz"), allow_pickle=False)
    nb = int(z["nb"])
    blocks = []
    for b in range(nb):
        blocks.append(cs.Block(tuple(int(x) for x in z[f"b{b}_dims"]), z[f"b{b}_nodes"], z[f"b{b}_T"],
                               z[f"b{b}_matID"], z[f"b{b}_qvol"]))
    params = cs.Params(**json.loads(bytes(z["params"]).decode()))
    bcs, sids = [], []
    for m in range(params.mgl):
        chim = {}
        if f"bc{m}_chent" in z:
            off = 0
            for e, n1, n2 in zip(z[f"bc{m}_chent"], z[f"bc{m}_chn1"], z[f"bc{m}_chn2"]):
                sl = slice(off, off + int(n1) + int(n2))
                chim[int(e)] = (int(n1), z[f"bc{m}_chdon"][sl], z[f"bc{m}_chvf"][sl])
                off = sl.stop
        bcs.append(fm.BCList(z[f"bc{m}_ib"], z[f"bc{m}_rb"], [None] * z[f"bc{m}_ib"].shape[0], chim))
        sids.append(z[f"bc{m}_sid"])
    series = [(int(z[f"s{s}_periodic"]), z[f"s{s}_t"], z[f"s{s}_v"]) for s in range(int(z["nseries"]))]
    tab = fm.PropertyTable(int(z["tab_Ti1"]), z["tab_cp"], z["tab_rho"], z["tab_kap"], z["tab_h"], [])
    c = cs.Case(blocks, bcs, tab, params, series, sids, name=name)
    return c, z
