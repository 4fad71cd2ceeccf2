"""Algorithmic (compulsory) HBM bytes of each kernel of the explicit step.

Counts every array element a kernel must read or write once per cell, at the
precision it is stored in; neighbour re-reads of the stencil and of the
face weights are excluded (they are served from cache/LDS), as are the
surface-only BC records.

FP64 storage: T, dt, qvol, vol, weights, R: 8 B; material index: 4 B.  A
level whose blocks each hold one material reads no material stream
(mat=False): the index is a per-block constant; likewise a block of uniform
qvol reads no qvol stream (qvol=False).  The four-stage update in primitive
variables does not read the stage state (stage_state=False).
"""
from __future__ import annotations


def stage_bytes(ndir=3, tang=False, irs=True, prim=True, fp=8, mat=True, qvol=True, rect=False, dt_stream=True):
    """Fused flux + BC + newstate kernel, per cell.

    reads: T, dt, qvol, vol, material, own face weights (ndir normal weights,
    or 3 per face when tangential weights are present); writes the scaled
    residual (IRS on) or reads T^n and writes T (IRS off).  rect: volume and
    weights are products of per-axis factors (no per-cell geometry read).
    dt_stream=False: a time-accurate step's global dt (one value, no stream)."""
    nw = 0 if rect else ndir * (3 if tang else 1)
    b = fp * (1 + (1 if dt_stream else 0) + (1 if qvol else 0) + (0 if rect else 1) + nw) + (4 if mat else 0)
    if irs:
        b += fp            # write R
    else:
        b += fp + fp       # read T^n, write T
    return b


def irs_bytes(update=False, fp=8, mat=True):
    """One IRS line pass: read R, write R; the last pass instead applies the
    RK update: read R, T, T^n, material, write T."""
    if update:
        return fp * 4 + (4 if mat else 0)
    return 2 * fp


def irs_fused_bytes(fp=8, mat=True, stage_state=True):
    """The x-y-z IRS pass with the update (levels without j/k connections):
    read R, T (the stage state, if the update uses it), T^n (and the
    material), write T."""
    return fp * (4 if stage_state else 3) + (4 if mat else 0)


def dt_bytes(fp=8, mat=True):
    """Local dt: read T, material, sum 1/dl^2, write dt."""
    return fp + (4 if mat else 0) + 8 + fp


def iteration_bytes(n_rk=3, ndir=3, tang=False, irs=True, prim=True, fp=8, mat=True, fused_irs=False, qvol=True,
                    rect=False, fused_dt=False):
    """Bytes per cell of one Explicit_Step.  fused_dt: the last stage's IRS
    pass computes the next step's dt from the new state (it writes dt and,
    unless rect, reads sum 1/dl^2) instead of a separate pass over T."""
    b = (fp + (0 if rect else 8)) if fused_dt else dt_bytes(fp, mat)
    per_stage = stage_bytes(ndir, tang, irs, prim, fp, mat, qvol, rect)
    if irs:
        if fused_irs:
            per_stage += irs_fused_bytes(fp, mat, stage_state=not (prim and n_rk == 4))
        else:
            per_stage += (ndir - 1) * irs_bytes(False, fp) + irs_bytes(True, fp, mat)
    return b + n_rk * per_stage


def cmg_stage_bytes(fp=4):
    """One stage of the correction smoother (no IRS): read d, d_0, b, the
    cell's three low-face conductances and the BC diagonal, write d."""
    return fp * 8


def canonical_bytes(n_rk=4, irs=True, ndir=3):
    """Per cell-update bytes of a straightforward FP64 implementation that
    streams every operand per stage: T, T^n, T written, nine face-geometry
    coefficients (three per owned face), 1/V, dt and qvol = 120 B per stage;
    T, three edge lengths read and dt written = 40 B per step; and with IRS
    16 B (R read and written) per direction and stage."""
    return n_rk * 120 + 40 + (ndir * 16 * n_rk if irs else 0)
