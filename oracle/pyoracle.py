"""ctypes wrapper of the CPU oracle (oracle/_build/libfuss_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product package.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "_build", "libfuss_oracle.so")
_LIB_OMP = os.path.join(_HERE, "_build", "libfuss_oracle_omp.so")
_lib = None
_lib_omp = None

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib(omp=False):
    """The scalar checker, or (omp=True) the OpenMP build of the same source
    used as the host-core CPU baseline."""
    global _lib, _lib_omp
    if (_lib_omp if omp else _lib) is None:
        path = _LIB_OMP if omp else _LIB
        if not os.path.isfile(path):
            build()
        L = C.CDLL(path)
        L.ob_create.restype = C.c_void_p
        L.ob_create.argtypes = [C.c_int, C.c_int]
        L.ob_destroy.argtypes = [C.c_void_p]
        L.ob_set_table.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, _dp, _dp, _dp, _dp]
        L.ob_set_block.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, _dp, _dp, _dp, _dp]
        L.ob_add_series.argtypes = [C.c_void_p, C.c_int, C.c_int, _dp, _dp]
        L.ob_set_bc.argtypes = [C.c_void_p, C.c_int, C.c_int, _ip, _dp, _ip]
        L.ob_set_params.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_int, C.c_int, C.c_int, C.c_double]
        L.ob_set_irs_mode.argtypes = [C.c_void_p, C.c_int]
        L.ob_setup.argtypes = [C.c_void_p, C.c_double]
        L.ob_step.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_double)]
        L.ob_restrict.argtypes = [C.c_void_p, C.c_int]
        L.ob_prolong.argtypes = [C.c_void_p, C.c_int]
        L.ob_run.argtypes = [C.c_void_p, _ip, C.c_double, C.c_double, C.c_int, C.c_int,
                             C.POINTER(C.c_int), _ip, _dp, _dp, C.c_int, _ip, _dp]
        L.ob_get.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_char_p, _dp]
        L.ob_set.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_char_p, _dp]
        L.ob_dims.argtypes = [C.c_void_p, C.c_int, C.c_int, _ip]
        L.ob_ndir.argtypes = [C.c_void_p]
        L.ob_time.restype = C.c_double
        L.ob_time.argtypes = [C.c_void_p, C.c_int]
        L.ob_dtglobal.restype = C.c_double
        L.ob_dtglobal.argtypes = [C.c_void_p, C.c_int]
        L.ob_last_error.argtypes = [C.c_void_p, _ip]
        L.ob_get_metrics.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_void_p,
                                     C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]
        L.ob_get_bc_ghost.argtypes = [C.c_void_p, C.c_int, C.c_int, _dp, _dp, _dp]
        L.ob_wall.argtypes = [C.c_void_p, C.c_int, _dp, _dp]
        L.ob_set_chimera.argtypes = [C.c_void_p, C.c_int, C.c_int, _ip, _ip, _ip, _ip, _dp]
        L.ob_threads.restype = C.c_int
        L.ob_threads.argtypes = []
        if omp:
            _lib_omp = L
        else:
            _lib = L
    return _lib_omp if omp else _lib


class OracleError(RuntimeError):
    pass


def _c(a, dt=np.float64):
    return np.ascontiguousarray(a, dtype=dt)


class Oracle:
    """One oracle context holding every multigrid level of a Case."""

    def __init__(self, case, params=None, omp=False):
        self.L = lib(omp)
        self.case = case
        p = params or case.params
        self.p = p
        self.mgl = p.mgl
        self.h = self.L.ob_create(p.mgl, case.nb)
        t = case.table
        self._chk(self.L.ob_set_table(self.h, t.nmat, t.Ti1, t.nT, _c(t.cp), _c(t.rho), _c(t.kap), _c(t.h)))
        for bi, b in enumerate(case.blocks):
            ni, nj, nk = b.dims
            self._chk(self.L.ob_set_block(self.h, bi + 1, ni, nj, nk, _c(b.nodes), _c(b.T), _c(b.matID),
                                          _c(b.qvol)))
        for s in case.series:
            self.L.ob_add_series(self.h, int(s[0]), len(s[1]), _c(s[1]), _c(s[2]))
        for m in range(p.mgl):
            bcl = case.bcs[m]
            sid = case.series_id[m] if m < len(case.series_id) else np.full(bcl.n, -1, np.int32)
            self._chk(self.L.ob_set_bc(self.h, m + 1, bcl.n, _c(bcl.ib, np.int32), _c(bcl.rb), _c(sid, np.int32)))
            if getattr(bcl, "chim", None):
                ent, n1, n2, don, vf = bcl.chimera_arrays()
                self._chk(self.L.ob_set_chimera(self.h, m + 1, len(ent), ent, n1, n2, _c(don.ravel(), np.int32),
                                                _c(vf)))
        self._chk(self.L.ob_set_params(self.h, p.n_rk, p.vnn, p.rampa_iter, int(p.time_accurate), int(p.prim),
                                       p.irs_beta))
        self._chk(self.L.ob_set_irs_mode(self.h, int(getattr(p, "irs_mode", 0))))
        self._chk(self.L.ob_setup(self.h, 0.0 if p.time_accurate else -1.0))

    def _chk(self, st):
        if st:
            bijk = np.zeros(4, np.int32)
            self.L.ob_last_error(self.h, bijk) if getattr(self, "h", None) else None
            raise OracleError(f"oracle status {st} at b,i,j,k={tuple(bijk)}")

    def __del__(self):
        try:
            if getattr(self, "h", None):
                self.L.ob_destroy(self.h)
                self.h = None
        except Exception:
            pass

    def dims(self, level, b):
        d = np.zeros(3, np.int32)
        self.L.ob_dims(self.h, level, b, d)
        return tuple(int(x) for x in d)

    def get(self, name, level=1, b=1):
        ni, nj, nk = self.dims(level, b)
        if name == "Tpad":
            out = np.zeros((nk + 4) * (nj + 4) * (ni + 4))
            self._chk(self.L.ob_get(self.h, level, b, b"Tpad", out))
            return out.reshape(nk + 4, nj + 4, ni + 4)
        out = np.zeros(ni * nj * nk)
        self._chk(self.L.ob_get(self.h, level, b, name.encode(), out))
        return out.reshape(nk, nj, ni)

    def set(self, name, arr, level=1, b=1):
        self._chk(self.L.ob_set(self.h, level, b, name.encode(), _c(arr)))

    def T(self, level=1):
        return [self.get("T", level, b + 1) for b in range(self.case.nb)]

    def step(self, level=1):
        r = C.c_double(0.0)
        self._chk(self.L.ob_step(self.h, level, C.byref(r)))
        return r.value

    def restrict(self, fine_level):
        self._chk(self.L.ob_restrict(self.h, fine_level))

    def prolong(self, fine_level):
        self._chk(self.L.ob_prolong(self.h, fine_level))

    def time(self, level=1):
        return self.L.ob_time(self.h, level)

    def dtglobal(self, level=1):
        return self.L.ob_dtglobal(self.h, level)

    def ndir(self):
        return self.L.ob_ndir(self.h)

    def metrics(self, level=1, b=1):
        ni, nj, nk = self.dims(level, b)
        npad = (ni + 4) * (nj + 4) * (nk + 4)
        M = np.zeros(9 * npad)
        dl = np.zeros(3 * npad)
        vol = np.zeros(npad)
        faces = []
        for d in range(3):
            e = [ni, nj, nk]
            e[d] += 1
            nf = e[0] * e[1] * e[2]
            fN = np.zeros(3 * nf)
            fA = np.zeros(nf)
            self.L.ob_get_metrics(self.h, level, b, None, None, None, d + 1,
                                  fN.ctypes.data_as(C.c_void_p), fA.ctypes.data_as(C.c_void_p))
            faces.append((fN.reshape(e[2], e[1], e[0], 3), fA.reshape(e[2], e[1], e[0])))
        self.L.ob_get_metrics(self.h, level, b, M.ctypes.data_as(C.c_void_p), dl.ctypes.data_as(C.c_void_p),
                              vol.ctypes.data_as(C.c_void_p), 0, None, None)
        sh = (nk + 4, nj + 4, ni + 4)
        return M.reshape(sh + (3, 3)).swapaxes(-1, -2), dl.reshape(sh + (3,)), vol.reshape(sh), faces

    def wall(self, level=1):
        """(Tw, qw) of every BC entry of the level at the current state
        (Extract_Properties_From_BC; 0 for entries that are not walls)."""
        n = int(self.case.bcs[level - 1].n)
        Tw, qw = np.zeros(n), np.zeros(n)
        self._chk(self.L.ob_wall(self.h, level, Tw, qw))
        return Tw, qw

    def bc_ghost(self, n, level=1):
        Mg = np.zeros(9)
        mat = np.zeros(1)
        Tg = np.zeros(6)
        self._chk(self.L.ob_get_bc_ghost(self.h, level, n, Mg, mat, Tg))
        return Mg.reshape(3, 3).T.copy(), float(mat[0]), Tg

    def run(self, max_iters=10**9, hist_cap=None, probes=None, itermax=None):
        """Run the FUSS_solve loop to convergence / limits; returns
        (n_iter, hist[iter_general, time, residual], probe_rows)."""
        p = self.p
        im = np.asarray(itermax if itermax is not None else p.itermax(), dtype=np.int32)
        cap = hist_cap if hist_cap is not None else min(max_iters, 10_000_000)
        hi = np.zeros(max(cap, 1), np.int32)
        ht = np.zeros(max(cap, 1))
        hr = np.zeros(max(cap, 1))
        npb = 0 if probes is None else len(probes)
        pr = _c(np.zeros((max(npb, 1), 4)), np.int32) if npb == 0 else _c(np.asarray(probes), np.int32)
        pout = np.zeros((max(cap, 1)) * (npb + 1))
        n = C.c_int(0)
        st = self.L.ob_run(self.h, im, p.res_threshold, p.time_threshold, int(min(max_iters, 2**31 - 1)), cap,
                           C.byref(n), hi, ht, hr, npb, pr, pout)
        self._chk(st)
        k = min(n.value, cap)
        hist = np.stack([hi[:k].astype(np.float64), ht[:k], hr[:k]], axis=1)
        prow = pout[: k * (npb + 1)].reshape(k, npb + 1) if npb else None
        return n.value, hist, prow
