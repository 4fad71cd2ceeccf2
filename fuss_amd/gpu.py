"""ctypes binding of libfuss_gpu.so and the FUSS driver loop on top of it.

``Solver`` mirrors the reference's library API ``FUSS_type%setup / solve``
(src/lib/driver/Procedures_m.f90:8-14): ``Solver(case)`` is FUSS_setup,
``explicit_step`` is Explicit_Step (src/lib/numerics/time/explicit/Mod_Explicit.f90:10-111)
and ``solve`` runs the FUSS_solve / FUSS_postprocess control of
src/lib/driver/Wrap_Solve.f90:10-33 and Wrap_Postprocess.f90:183-187
(coarse-to-fine multigrid sequencing, residual/time/iteration stops).

Errors are raised as ``FussError`` carrying the C-ABI status and, for a failed
state check, the (block, i, j, k) of the cell -- the information the reference
prints before ``stop "NaN or T<0 detected"`` (Lib_Newstate.f90:283-287).
There is no CPU fallback: a missing library is an ImportError.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libfuss_gpu.so")

FUSS_FP64, FUSS_FP32 = 0, 1
FUSS_CONS, FUSS_PRIM = 0, 1
FUSS_IRS_JACOBI2, FUSS_IRS_TRIDIAG = 0, 1
FUSS_ESTATE = 2
FUSS_EARG = 1
FUSS_ETABLE = 3
FUSS_EMETRIC = 4
FUSS_EUNSUPPORTED = 5
FUSS_EBC = 7
FUSS_STAGE_REGCOL = 1  # fuss_set_tiles stage_rows: the register-column stage kernel

# fuss_stage_fn: (user, level, irk, block, ni, nj, nk, const double *T, double *R)
STAGE_FN = C.CFUNCTYPE(None, C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                       C.POINTER(C.c_double), C.POINTER(C.c_double))


class StepParams(C.Structure):
    _fields_ = [("n_rk", C.c_int), ("vnn", C.c_double), ("rampa_iter", C.c_int), ("time_accurate", C.c_int),
                ("integration", C.c_int), ("irs_beta", C.c_double), ("irs_mode", C.c_int),
                ("compute_residual", C.c_int)]


class StepResult(C.Structure):
    _fields_ = [("res_sq", C.c_double), ("dtglobal", C.c_double), ("time", C.c_double), ("iter", C.c_int),
                ("status", C.c_int), ("bijk", C.c_int * 4)]


class RefineParams(C.Structure):
    _fields_ = [("precision", C.c_int), ("n_stage", C.c_int), ("lam_lo", C.c_double), ("lam_hi", C.c_double),
                ("irs_beta", C.c_double), ("nu1", C.c_int), ("nu2", C.c_int), ("vcycles", C.c_int),
                ("max_levels", C.c_int), ("ncoarse", C.c_int), ("max_outer", C.c_int), ("res_threshold", C.c_double),
                ("per_block", C.c_int), ("coarse_cg", C.c_int)]


class RefineResult(C.Structure):
    _fields_ = [("outer", C.c_int), ("vcycles", C.c_int), ("levels", C.c_int), ("converged", C.c_int),
                ("res", C.c_double), ("defect", C.c_double)]


def refine_params(precision=FUSS_FP32, n_stage=4, lam_lo=0.3, lam_hi=2.0, irs_beta=0.0, nu1=1, nu2=1, vcycles=2,
                  max_levels=0, ncoarse=20, max_outer=200, res_threshold=1e-10, per_block=False, coarse_cg=0):
    """Defaults: FP32 sweeps, four-stage Chebyshev smoother on [0.3, 2], one
    pre- and one post-smoothing iteration, two V-cycles per outer iteration,
    coarsening down to <= 2 cells per direction.  coarse_cg: conjugate-
    gradient iterations on a coarsest level larger than the direct solve
    (with max_levels > 0), 0: ncoarse smoothing iterations."""
    return RefineParams(precision, n_stage, lam_lo, lam_hi, irs_beta, nu1, nu2, vcycles, max_levels, ncoarse,
                        max_outer, res_threshold, int(per_block), int(coarse_cg))


_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_lib = None


def load():
    """Load libfuss_gpu.so; raise ImportError when it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.isfile(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: run __graft_entry__.build() (hipcc, gfx950)")
    L = C.CDLL(LIB_PATH)
    P = C.c_void_p
    L.fuss_create.argtypes = [C.c_int, C.c_int, C.c_int, C.POINTER(P)]
    L.fuss_set_table.argtypes = [P, C.c_int, C.c_int, C.c_int, _dp, _dp, _dp, _dp]
    L.fuss_set_block.argtypes = [P, C.c_int, C.c_int, C.c_int, C.c_int, _dp, _dp, _dp, _dp]
    L.fuss_add_series.argtypes = [P, C.c_int, C.c_int, _dp, _dp, C.POINTER(C.c_int)]
    L.fuss_set_bc.argtypes = [P, C.c_int, C.c_int, _ip, _dp, _ip]
    L.fuss_set_chimera.argtypes = [P, C.c_int, C.c_int, _ip, _ip, _ip, _ip, _dp]
    L.fuss_setup.argtypes = [P, C.c_double]
    L.fuss_step.argtypes = [P, C.c_int, C.POINTER(StepParams), C.POINTER(StepResult)]
    L.fuss_run.argtypes = [P, C.c_int, C.POINTER(StepParams), C.c_int, C.c_int, C.c_double, C.c_double,
                           C.POINTER(C.c_int), C.c_void_p, C.c_void_p, C.POINTER(StepResult)]
    L.fuss_restrict.argtypes = [P, C.c_int]
    L.fuss_prolong.argtypes = [P, C.c_int]
    L.fuss_get_T.argtypes = [P, C.c_int, C.c_int, _dp]
    L.fuss_set_T.argtypes = [P, C.c_int, C.c_int, _dp]
    L.fuss_probe.argtypes = [P, C.c_int, _ip, _dp]
    L.fuss_get_clock.argtypes = [P, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_int)]
    L.fuss_set_clock.argtypes = [P, C.c_int, C.c_double, C.c_int]
    L.fuss_ndir.argtypes = [P, C.POINTER(C.c_int)]
    L.fuss_get_field.argtypes = [P, C.c_int, C.c_int, C.c_char_p, _dp]
    L.fuss_get_metrics.argtypes = [P, C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]
    L.fuss_get_stream.argtypes = [P, C.POINTER(P)]
    L.fuss_set_stream.argtypes = [P, P]
    L.fuss_memory.argtypes = [P, C.POINTER(C.c_double)]
    L.fuss_rect.argtypes = [P, C.c_int]
    L.fuss_wall.argtypes = [P, C.c_int, _dp, _dp]
    L.fuss_step_hook.argtypes = [P, C.c_int, C.POINTER(StepParams), STAGE_FN, C.c_void_p, C.POINTER(StepResult)]
    L.fuss_rect_info.argtypes = [P, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_double)]
    L.fuss_graphs.argtypes = [P, C.c_int]
    L.fuss_persistent.argtypes = [P, C.c_int]
    L.fuss_fused.argtypes = [P, C.c_int]
    L.fuss_autotune.argtypes = [P, C.c_int, C.POINTER(StepParams), C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                C.POINTER(C.c_double)]
    L.fuss_set_tiles.argtypes = [P, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]
    L.fuss_tiles.argtypes = [P, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int),
                             C.POINTER(C.c_int)]
    L.fuss_refine_setup.argtypes = [P, C.POINTER(RefineParams)]
    L.fuss_refine.argtypes = [P, C.POINTER(StepParams), C.POINTER(RefineResult), C.c_void_p]
    L.fuss_refine_blocks.argtypes = [P, C.c_void_p, C.c_void_p]
    L.fuss_strerror.restype = C.c_char_p
    L.fuss_strerror.argtypes = [C.c_int]
    L.fuss_destroy.argtypes = [P]
    _lib = L
    return L


class FussError(RuntimeError):
    def __init__(self, code, what="", bijk=None):
        self.code = code
        self.bijk = tuple(bijk) if bijk is not None else None
        msg = f"fuss_gpu: {load().fuss_strerror(code).decode()} (status {code})"
        if what:
            msg += f" in {what}"
        if bijk is not None and code == FUSS_ESTATE:
            msg += f"; Integration failed at b, i, j, k: {tuple(bijk)}"
        super().__init__(msg)


def _chk(code, what=""):
    if code:
        raise FussError(code, what)


def _c(a, dt=np.float64):
    return np.ascontiguousarray(a, dtype=dt)


def step_params(p, compute_residual=True, irs_mode=None, n_rk=None):
    """StepParams of the case numerics p; irs_mode None: p.irs_mode."""
    mode = getattr(p, "irs_mode", FUSS_IRS_JACOBI2) if irs_mode is None else irs_mode
    return StepParams(n_rk or p.n_rk, p.vnn, p.rampa_iter, int(p.time_accurate),
                      FUSS_PRIM if p.prim else FUSS_CONS, p.irs_beta, mode, int(compute_residual))


class Solver:
    """A case resident on the GPU: FUSS_setup done at construction."""

    def __init__(self, case, precision=FUSS_FP64, params=None):
        self.L = load()
        self.case = case
        self.p = params or case.params
        self.h = C.c_void_p()
        _chk(self.L.fuss_create(self.p.mgl, case.nb, precision, C.byref(self.h)), "fuss_create")
        t = case.table
        _chk(self.L.fuss_set_table(self.h, t.nmat, t.Ti1, t.nT, _c(t.cp), _c(t.rho), _c(t.kap), _c(t.h)),
             "fuss_set_table")
        for bi, b in enumerate(case.blocks):
            ni, nj, nk = b.dims
            _chk(self.L.fuss_set_block(self.h, bi + 1, ni, nj, nk, _c(b.nodes), _c(b.T), _c(b.matID), _c(b.qvol)),
                 "fuss_set_block")
        for s in case.series:
            sid = C.c_int(0)
            _chk(self.L.fuss_add_series(self.h, int(s[0]), len(s[1]), _c(s[1]), _c(s[2]), C.byref(sid)))
        for m in range(self.p.mgl):
            bcl = case.bcs[m]
            sid = case.series_id[m] if m < len(case.series_id) else np.full(bcl.n, -1, np.int32)
            _chk(self.L.fuss_set_bc(self.h, m + 1, bcl.n, _c(bcl.ib, np.int32), _c(bcl.rb), _c(sid, np.int32)),
                 f"fuss_set_bc level {m+1}")
            if getattr(bcl, "chim", None):
                ent, n1, n2, don, vf = bcl.chimera_arrays()
                _chk(self.L.fuss_set_chimera(self.h, m + 1, len(ent), ent, n1, n2, _c(don.ravel(), np.int32), _c(vf)),
                     f"fuss_set_chimera level {m+1}")
        _chk(self.L.fuss_setup(self.h, 0.0 if self.p.time_accurate else -1.0), "fuss_setup")
        self.level = self.p.mgl

    def autotune(self, level=1, steps=5):
        """Choose the column height and stage kernel of a 3-D level by timing
        (fuss_autotune); returns (kc, row_march, ms per step)."""
        sp = step_params(self.p)
        kc, row, ms = C.c_int(0), C.c_int(0), C.c_double(0)
        _chk(self.L.fuss_autotune(self.h, level, C.byref(sp), int(steps), C.byref(kc), C.byref(row), C.byref(ms)),
             "fuss_autotune")
        return kc.value, row.value, ms.value

    def set_tiles(self, level=1, kc=8, row_march=-1, stage_rows=4, irs_rows=4):
        """Column tiles of a 3-D level without timing (fuss_set_tiles)."""
        _chk(self.L.fuss_set_tiles(self.h, level, int(kc), int(row_march), int(stage_rows), int(irs_rows)),
             "fuss_set_tiles")

    def tiles(self, level=1):
        """(kc, row_march, stage_rows, irs_rows) of a level (fuss_tiles; 0s on 2-D levels)."""
        kc, row, st, ty = C.c_int(0), C.c_int(0), C.c_int(0), C.c_int(0)
        _chk(self.L.fuss_tiles(self.h, level, C.byref(kc), C.byref(row), C.byref(st), C.byref(ty)), "fuss_tiles")
        return kc.value, row.value, st.value, ty.value

    def persistent(self, enable):
        """fuss_run as one persistent launch on eligible levels (fuss_persistent)."""
        _chk(self.L.fuss_persistent(self.h, int(bool(enable))), "fuss_persistent")

    def fused(self, enable):
        """Fused stages on 2-D levels without IRS and connections (fuss_fused)."""
        _chk(self.L.fuss_fused(self.h, int(bool(enable))), "fuss_fused")

    def graphs(self, enable):
        """fuss_run on captured HIP graphs (default) or step by step."""
        _chk(self.L.fuss_graphs(self.h, int(bool(enable))))

    def wall(self, level=1):
        """(Tw, qw) of every BC entry of the level at the current state, in
        BC-list order (fuss_wall; 0 for entries that are not walls)."""
        n = int(self.case.bcs[level - 1].n)
        Tw, qw = np.zeros(n), np.zeros(n)
        _chk(self.L.fuss_wall(self.h, level, Tw, qw), "fuss_wall")
        return Tw, qw

    def rect(self, enable):
        """Per-axis metric factors on rectilinear levels (default) or the
        stored 3-D metric arrays (fuss_rect)."""
        _chk(self.L.fuss_rect(self.h, int(bool(enable))), "fuss_rect")

    def rect_info(self, level=1):
        """(weights from factors, dt from factors, largest relative deviation
        of the factors from the stored metrics; -1 when not rectilinear)."""
        w, d, dev = C.c_int(0), C.c_int(0), C.c_double(0)
        _chk(self.L.fuss_rect_info(self.h, level, C.byref(w), C.byref(d), C.byref(dev)), "fuss_rect_info")
        return bool(w.value), bool(d.value), dev.value

    def close(self):
        if getattr(self, "h", None) and self.h.value:
            self.L.fuss_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- state access ------------------------------------------------------
    def dims(self, level, b):
        ni, nj, nk = self.case.blocks[b - 1].dims
        for _ in range(1, level):
            ni, nj, nk = ni // 2, nj // 2, max(1, nk // 2)
        return ni, nj, nk

    def get(self, name, level=1, b=1):
        ni, nj, nk = self.dims(level, b)
        out = np.zeros(ni * nj * nk)
        _chk(self.L.fuss_get_field(self.h, level, b, name.encode(), out), f"fuss_get_field {name}")
        return out.reshape(nk, nj, ni)

    def T(self, level=1):
        return [self.get("T", level, b + 1) for b in range(self.case.nb)]

    def set_T(self, T, level=1, b=1):
        _chk(self.L.fuss_set_T(self.h, level, b, _c(T)), "fuss_set_T")

    def set_T_all(self, Ts, level=1):
        """T of every block of the level (a list in block order)."""
        for b, T in enumerate(Ts):
            self.set_T(T, level, b + 1)

    def set_clock(self, level, time, iteration):
        """Level time and iteration counter (restart; fuss_set_clock)."""
        _chk(self.L.fuss_set_clock(self.h, level, float(time), int(iteration)), "fuss_set_clock")

    def probe(self, bijk):
        q = _c(np.asarray(bijk).reshape(-1, 4), np.int32)
        out = np.zeros(q.shape[0])
        _chk(self.L.fuss_probe(self.h, q.shape[0], q, out), "fuss_probe")
        return out

    def clock(self, level=1):
        t = C.c_double(0)
        it = C.c_int(0)
        _chk(self.L.fuss_get_clock(self.h, level, C.byref(t), C.byref(it)))
        return t.value, it.value

    def metrics(self, level=1, b=1):
        ni, nj, nk = self.dims(level, b)
        M = np.zeros(9 * ni * nj * nk)
        vol = np.zeros(ni * nj * nk)
        _chk(self.L.fuss_get_metrics(self.h, level, b, M.ctypes.data_as(C.c_void_p), vol.ctypes.data_as(C.c_void_p),
                                     0, None, None))
        faces = []
        for d in range(3):
            e = [ni, nj, nk]
            e[d] += 1
            N = np.zeros(3 * e[0] * e[1] * e[2])
            A = np.zeros(e[0] * e[1] * e[2])
            _chk(self.L.fuss_get_metrics(self.h, level, b, None, None, d + 1, N.ctypes.data_as(C.c_void_p),
                                         A.ctypes.data_as(C.c_void_p)))
            faces.append((N.reshape(e[2], e[1], e[0], 3), A.reshape(e[2], e[1], e[0])))
        return M.reshape(nk, nj, ni, 3, 3).swapaxes(-1, -2), vol.reshape(nk, nj, ni), faces

    def stream(self):
        s = C.c_void_p()
        _chk(self.L.fuss_get_stream(self.h, C.byref(s)))
        return s.value

    # -- hot path ----------------------------------------------------------
    def explicit_step(self, level=None, compute_residual=True, irs_mode=None, hook=None):
        """One Explicit_Step; returns the StepResult (res_sq, dtglobal, time, iter).

        hook(level, irk, block, T, R): the External_Function slot, called per
        stage and block with the stage input T and the raw residual R as
        (nk, nj, ni) arrays over host memory; modify R in place (fuss_step_hook)."""
        lv = level or self.level
        sp = step_params(self.p, compute_residual, irs_mode)
        r = StepResult()
        if hook is None:
            code = self.L.fuss_step(self.h, lv, C.byref(sp), C.byref(r))
        else:
            def cb(user, level, irk, block, ni, nj, nk, Tp, Rp):
                shape = (nk, nj, ni)
                hook(level, irk, block, np.ctypeslib.as_array(Tp, shape=shape),
                     np.ctypeslib.as_array(Rp, shape=shape))
            fn = STAGE_FN(cb)
            code = self.L.fuss_step_hook(self.h, lv, C.byref(sp), fn, None, C.byref(r))
        if code:
            raise FussError(code, "fuss_step", list(r.bijk))
        return r

    def restrict(self, fine_level):
        _chk(self.L.fuss_restrict(self.h, fine_level), "fuss_restrict")

    def prolong(self, fine_level):
        _chk(self.L.fuss_prolong(self.h, fine_level), "fuss_prolong")

    def run_level(self, level, max_iters, itermax_left, res_threshold, time_threshold, hist=True,
                  irs_mode=None):
        sp = step_params(self.p, True, irs_mode)
        n = C.c_int(0)
        hr = np.zeros(max(max_iters, 1)) if hist else None
        ht = np.zeros(max(max_iters, 1)) if hist else None
        r = StepResult()
        code = self.L.fuss_run(self.h, level, C.byref(sp), int(max_iters), int(itermax_left), res_threshold,
                               time_threshold, C.byref(n), hr.ctypes.data_as(C.c_void_p) if hist else None,
                               ht.ctypes.data_as(C.c_void_p) if hist else None, C.byref(r))
        if code:
            raise FussError(code, "fuss_run", list(r.bijk))
        k = n.value
        return k, (hr[:k] if hist else None), (ht[:k] if hist else None), r

    def solve(self, max_iters=10**9, irs_mode=None, chunk=None, progress=None):
        """FUSS_solve loop to the end of the run: coarse-to-fine sequencing,
        then level 1 until a stop condition.  Returns (n_iter, history) with
        history rows (iter_general, time, residual).  chunk: at most this
        many steps per fuss_run call (the same steps and stop tests), after
        each of which progress(n_iter, level, residual) is called."""
        p = self.p
        itermax = p.itermax()
        rows = []
        done = 0
        while done < max_iters:
            lv = self.level
            lvdone = 0
            while True:
                n = max_iters - done if not chunk else min(chunk, max_iters - done)
                k, hr, ht, _ = self.run_level(lv, n, itermax[lv - 1] - lvdone, p.res_threshold,
                                              p.time_threshold, True, irs_mode)
                rows.append(np.stack([done + 1 + np.arange(k, dtype=np.float64), ht, hr], axis=1))
                done += k
                lvdone += k
                stop = (lvdone >= itermax[lv - 1] or (k and hr[-1] <= p.res_threshold)
                        or (k and ht[-1] >= p.time_threshold))
                if progress:
                    progress(done, lv, float(hr[-1]) if k else float("nan"))
                if stop or k < n or done >= max_iters:
                    break
            if lv == 1 or not stop or done >= max_iters:
                break
            self.prolong(lv - 1)
            self.level = lv - 1
        return done, (np.concatenate(rows) if rows else np.zeros((0, 3)))

    # -- steady state by defect correction --------------------------------
    def refine(self, rp=None, monitor=None, setup=True):
        """Steady state of the finest level by FP64 defect correction
        (fuss_refine).  monitor: StepParams of the FP64 monitor step (default:
        the case's numerics).  Returns (RefineResult, residual history)."""
        rp = rp or refine_params(res_threshold=self.p.res_threshold)
        if setup:
            _chk(self.L.fuss_refine_setup(self.h, C.byref(rp)), "fuss_refine_setup")
        mon = monitor or step_params(self.p)
        hist = np.zeros(rp.max_outer + 1)
        r = RefineResult()
        _chk(self.L.fuss_refine(self.h, C.byref(mon), C.byref(r), hist.ctypes.data_as(C.c_void_p)), "fuss_refine")
        self.level = 1
        return r, hist[: r.outer + 1]

    def refine_blocks(self):
        """Per block: (first outer iteration meeting the threshold or -1, last residual)."""
        it = np.zeros(self.case.nb, np.int32)
        res = np.zeros(self.case.nb)
        _chk(self.L.fuss_refine_blocks(self.h, it.ctypes.data_as(C.c_void_p), res.ctypes.data_as(C.c_void_p)))
        return it, res
