"""Algorithmic (compulsory) HBM bytes of each kernel of the explicit step.

Counts every array element a kernel must read or write once per cell, at the
precision it is stored in; neighbour re-reads of the stencil and of the
face weights are excluded (they are served from cache/LDS), as are the
surface-only BC records.

FP64 storage: T, dt, qvol, vol, weights, R: 8 B; material index: 4 B.
"""
from __future__ import annotations


def stage_bytes(ndir=3, tang=False, irs=True, prim=True, fp=8):
    """Fused flux + BC + newstate kernel, per cell.

    reads: T, dt, qvol, vol, material, own face weights (ndir normal weights,
    or 3 per face when tangential weights are present); writes the scaled
    residual (IRS on) or reads T^n and writes T (IRS off)."""
    nw = ndir * (3 if tang else 1)
    b = fp * (1 + 1 + 1 + 1 + nw) + 4
    if irs:
        b += fp            # write R
    else:
        b += fp + fp       # read T^n, write T
    return b


def irs_bytes(update=False, fp=8):
    """One IRS line pass: read R, write R; the last pass instead applies the
    RK update: read R, T, T^n, material, write T."""
    if update:
        return fp * 4 + 4
    return 2 * fp


def dt_bytes(fp=8):
    """Local dt: read T, material, sum 1/dl^2, write dt."""
    return fp + 4 + 8 + fp


def iteration_bytes(n_rk=3, ndir=3, tang=False, irs=True, prim=True, fp=8):
    b = dt_bytes(fp)
    per_stage = stage_bytes(ndir, tang, irs, prim, fp)
    if irs:
        per_stage += (ndir - 1) * irs_bytes(False, fp) + irs_bytes(True, fp)
    return b + n_rk * per_stage


def cmg_stage_bytes(fp=4):
    """One stage of the correction smoother (no IRS): read d, d_0, b, the
    cell's three low-face conductances and the BC diagonal, write d."""
    return fp * 8
