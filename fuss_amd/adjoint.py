"""Heat-flux / conductivity inverse problem on the GPU: a torch.autograd.Function
over the hand-written discrete adjoint of the transient explicit step.

The forward model is the reference's time-accurate Explicit_Step
(src/lib/numerics/time/explicit/Mod_Explicit.f90:10-111, global dt from
Set_Global_dt, Mod_dt.f90:86-107) run for ``nsteps`` steps on the device; its
outputs are T at probe cells after every step (IO_Probes.f90:124-166).  The
parameters are the 301 wall flux of each segment of a face (the BC value is
-q on a min face, Lib_BC_Fluxes_Wall_Heat.f90:26,44-47) and a scale c of the
conductivity table.  ``ProbeHistory.apply(q, c, problem)`` returns the
(nsteps, nprobe) history; its backward runs fuss_adj_backward (checkpointed
recompute + reverse sweep, fuss_amd/csrc/fuss_adjoint.h) through the same
C-ABI.  Only one forward's checkpoints are kept per problem: the backward of
an older forward raises.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import gpu


def _load():
    L = gpu.load()
    P = C.c_void_p
    L.fuss_set_kscale.argtypes = [P, C.c_double]
    L.fuss_set_bc_values.argtypes = [P, C.c_int, C.c_int, gpu._ip, C.c_int, gpu._dp]
    L.fuss_adj_setup.argtypes = [P, C.c_int, C.c_int, C.c_int, gpu._ip, C.c_int]
    L.fuss_adj_forward.argtypes = [P, C.POINTER(gpu.StepParams), C.c_void_p]
    L.fuss_adj_backward.argtypes = [P, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    L.fuss_set_clock.argtypes = [P, C.c_int, C.c_double, C.c_int]
    return L


class FluxProblem:
    """One GPU context holding the transient plate, its probes and checkpoints.

    seg_entries: (nseg, m) 0-based BC entries of level 1 sharing one flux
    value; probes: (nprobe, 4) 1-based (b, i, j, k)."""

    def __init__(self, case, seg_entries, probes, nsteps, ckpt_every=20, precision=gpu.FUSS_FP64):
        self.L = _load()
        self.solver = gpu.Solver(case)
        self.h = self.solver.h
        self.T0 = [np.ascontiguousarray(b.T, dtype=np.float64) for b in case.blocks]
        self.seg = np.ascontiguousarray(seg_entries, dtype=np.int32)
        self.entries = np.ascontiguousarray(self.seg.reshape(-1))
        self.nsteps = int(nsteps)
        pr = np.ascontiguousarray(np.asarray(probes, np.int32).reshape(-1, 4))
        self.nprobe = pr.shape[0]
        self.nbound = case.bcs[0].n
        self.sp = gpu.step_params(case.params, compute_residual=False)
        gpu._chk(self.L.fuss_adj_setup(self.h, self.nsteps, int(ckpt_every), self.nprobe, pr.reshape(-1),
                                       int(precision)), "fuss_adj_setup")
        self.generation = 0

    def close(self):
        self.solver.close()

    def set_params(self, q, kscale):
        q = np.asarray(q, np.float64).reshape(-1)
        vals = np.ascontiguousarray(np.repeat(-q, self.seg.shape[1]))
        gpu._chk(self.L.fuss_set_bc_values(self.h, 1, self.entries.size, self.entries, 0, vals), "fuss_set_bc_values")
        gpu._chk(self.L.fuss_set_kscale(self.h, float(kscale)), "fuss_set_kscale")

    def forward(self, q, kscale):
        """Probe history (nsteps, nprobe) on the GPU (float64) from the initial state."""
        self.set_params(q, kscale)
        for b, T in enumerate(self.T0):
            self.solver.set_T(T, 1, b + 1)
        gpu._chk(self.L.fuss_set_clock(self.h, 1, 0.0, 0))
        out = torch.empty((self.nsteps, self.nprobe), dtype=torch.float64, device="cuda")
        torch.cuda.synchronize()
        gpu._chk(self.L.fuss_adj_forward(self.h, C.byref(self.sp), C.c_void_p(out.data_ptr())), "fuss_adj_forward")
        self.generation += 1
        return out

    def backward(self, g, grad_T0=False):
        """dL/dq per segment (numpy), dL/dc (float) [, dL/dT^0] for the
        probe-history adjoint g (nsteps, nprobe) of the latest forward."""
        g = g.detach().to(device="cuda", dtype=torch.float64).contiguous()
        gq = torch.zeros(self.nbound, dtype=torch.float64, device="cuda")
        gk = torch.zeros(1, dtype=torch.float64, device="cuda")
        gT = torch.zeros(sum(T.size for T in self.T0), dtype=torch.float64, device="cuda") if grad_T0 else None
        torch.cuda.synchronize()
        gpu._chk(self.L.fuss_adj_backward(self.h, C.c_void_p(g.data_ptr()), C.c_void_p(gq.data_ptr()),
                                          C.c_void_p(gk.data_ptr()),
                                          C.c_void_p(gT.data_ptr()) if gT is not None else None),
                 "fuss_adj_backward")
        gqe = gq.cpu().numpy()[self.entries].reshape(self.seg.shape)
        out = (-gqe.sum(axis=1), float(gk.item()))
        if grad_T0:
            out = out + (gT.cpu().numpy(),)
        return out


class ProbeHistory(torch.autograd.Function):
    """probes = ProbeHistory.apply(q, kscale, problem): q (nseg,) and kscale
    (0-d) float64 tensors; returns the (nsteps, nprobe) probe history."""

    @staticmethod
    def forward(ctx, q, kscale, problem):
        out = problem.forward(q.detach().cpu().numpy(), float(kscale.detach().cpu()))
        ctx.problem = problem
        ctx.generation = problem.generation
        ctx.qdev, ctx.kdev = q.device, kscale.device
        return out

    @staticmethod
    def backward(ctx, g):
        pb = ctx.problem
        if pb.generation != ctx.generation:
            raise RuntimeError("ProbeHistory: a later forward replaced this forward's checkpoints")
        gq, gk = pb.backward(g)
        return (torch.as_tensor(gq, dtype=torch.float64, device=ctx.qdev),
                torch.tensor(gk, dtype=torch.float64, device=ctx.kdev), None)
