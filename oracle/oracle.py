"""ctypes binding of the CPU oracle (oracle/smaug_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
the cpu_baseline leg of bench.py, as the checker -- never by the product.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_build", "liboracle.so")


class SmGrid(ctypes.Structure):
    _fields_ = [
        ("ndim", ctypes.c_int),
        ("n", ctypes.c_int * 3),
        ("bc", ctypes.c_int * 3),
        ("hyperdif", ctypes.c_int),
        ("driver", ctypes.c_int),
        ("gamma", ctypes.c_double),
        ("g", ctypes.c_double * 3),
        ("chyp", ctypes.c_double * 13),
        ("drv", ctypes.c_double * 8),
        ("dx", ctypes.POINTER(ctypes.c_double) * 3),
        ("pos", ctypes.POINTER(ctypes.c_double) * 3),
        ("wfix", ctypes.POINTER(ctypes.c_double)),
    ]


_libs = {}


def build():
    subprocess.check_call(["make", "-s", "-C", HERE])


def lib(path=None):
    """the oracle library; ``path`` names another build of the same source
    (the mutation tests load copies with a planted mistake)"""
    path = LIB if path is None else path
    if path not in _libs:
        if path == LIB and not os.path.exists(LIB):
            build()
        L = ctypes.CDLL(path)
        dp = ctypes.POINTER(ctypes.c_double)
        gp = ctypes.POINTER(SmGrid)
        L.sm_ncell.argtypes = [gp]
        L.sm_ncell.restype = ctypes.c_long
        L.sm_boundary.argtypes = [gp, dp]
        L.sm_cmax.argtypes = [gp, dp, dp, dp]
        L.sm_maxcourant.argtypes = [gp, dp, dp]
        L.sm_maxcourant.restype = ctypes.c_double
        L.sm_maxcourant_dim.argtypes = [gp, dp, dp, ctypes.c_int]
        L.sm_maxcourant_dim.restype = ctypes.c_double
        L.sm_numax.argtypes = [gp, dp, dp, dp]
        L.sm_numax.restype = ctypes.c_double
        L.sm_dtvisc.argtypes = [gp, ctypes.c_double]
        L.sm_dtvisc.restype = ctypes.c_double
        L.sm_nu.argtypes = [gp, dp, dp, ctypes.c_int, ctypes.c_int, ctypes.c_double, dp, dp]
        L.sm_grad4.argtypes = [gp, dp, ctypes.c_int, dp]
        L.sm_rhs.argtypes = [gp, dp, dp, ctypes.c_double, dp, dp]
        L.sm_step_euler.argtypes = [gp, dp, dp, ctypes.c_double, ctypes.c_double, dp]
        L.sm_step_rk4.argtypes = [gp, dp, dp, ctypes.c_double, ctypes.c_double, dp]
        L.sm_run.argtypes = [gp, dp, dp, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                             ctypes.c_int]
        L.sm_run.restype = ctypes.c_double
        L.sm_num_threads.restype = ctypes.c_int
        _libs[path] = L
    return _libs[path]


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class Oracle:
    """Oracle bound to one set-up's grid (configs.Setup-like object)."""

    def __init__(self, s, lib_path=None):
        self.s = s
        self.L = lib(lib_path)
        g = SmGrid()
        g.ndim = s.ndim
        self._keep = []
        for d in range(3):
            g.n[d] = s.n[d] if d < len(s.n) else 1
            g.bc[d] = s.bc[d] if d < s.ndim else 0
            g.g[d] = s.g[d]
        g.hyperdif = s.hyperdif
        g.driver = s.driver
        g.gamma = s.gamma
        for f in range(13):
            g.chyp[f] = float(s.chyp[f])
        for k in range(8):
            g.drv[k] = float(s.drv[k])
        for d in range(s.ndim):
            dx = np.ascontiguousarray(s.dx[d], dtype=np.float64)
            pos = np.ascontiguousarray(s.pos[d], dtype=np.float64)
            self._keep += [dx, pos]
            g.dx[d] = _dp(dx)
            g.pos[d] = _dp(pos)
        self.g = g
        if getattr(s, "wfix", None) is not None:
            self.set_fixed(s.wfix)

    def set_fixed(self, w):
        """the state whose ghost values the fixed (bc 5) faces keep"""
        wf = np.ascontiguousarray(w, dtype=np.float64).copy()
        assert wf.shape == (self.s.nve, self.s.ncell)
        self._wfix = wf
        self.g.wfix = _dp(wf)

    def _w(self, w):
        w = np.ascontiguousarray(w, dtype=np.float64)
        assert w.shape == (self.s.nve, self.s.ncell)
        return w

    def boundary(self, w):
        w = self._w(w).copy()
        self.L.sm_boundary(ctypes.byref(self.g), _dp(w))
        return w

    def cmax(self, w, bg=None):
        bg = np.ascontiguousarray(self.s.bg if bg is None else bg)
        w = self._w(w)
        c = np.zeros(3)
        self.L.sm_cmax(ctypes.byref(self.g), _dp(w), _dp(bg), _dp(c))
        return c

    def maxcourant(self, w, bg=None):
        bg = np.ascontiguousarray(self.s.bg if bg is None else bg)
        w = self._w(w)
        return self.L.sm_maxcourant(ctypes.byref(self.g), _dp(w), _dp(bg))

    def maxcourant_dim(self, w, d, bg=None):
        bg = np.ascontiguousarray(self.s.bg if bg is None else bg)
        w = self._w(w)
        return self.L.sm_maxcourant_dim(ctypes.byref(self.g), _dp(w), _dp(bg), d)

    def numax(self, w, cmax=None, bg=None):
        bg = np.ascontiguousarray(self.s.bg if bg is None else bg)
        w = self._w(w)
        cp = None
        if cmax is not None:
            cm = np.ascontiguousarray(cmax, dtype=np.float64)
            self._keep.append(cm)
            cp = _dp(cm)
        return self.L.sm_numax(ctypes.byref(self.g), _dp(w), _dp(bg), cp)

    def dtvisc(self, numax):
        return self.L.sm_dtvisc(ctypes.byref(self.g), float(numax))

    def nu(self, w, field, dim, cmax_d, bg=None):
        bg = np.ascontiguousarray(self.s.bg if bg is None else bg)
        w = self._w(w)
        nr = np.zeros(self.s.ncell)
        nl = np.zeros(self.s.ncell)
        self.L.sm_nu(ctypes.byref(self.g), _dp(w), _dp(bg), field, dim, cmax_d, _dp(nr), _dp(nl))
        return nr, nl

    def grad4(self, q, d):
        q = np.ascontiguousarray(q, dtype=np.float64)
        out = np.zeros(self.s.ncell)
        self.L.sm_grad4(ctypes.byref(self.g), _dp(q), d, _dp(out))
        return out

    def rhs(self, w, t=0.0, bg=None, cmax=None):
        bg = np.ascontiguousarray(self.s.bg if bg is None else bg)
        w = self._w(w)
        r = np.zeros_like(w)
        cp = None
        if cmax is not None:
            cm = np.ascontiguousarray(cmax, dtype=np.float64)
            self._keep.append(cm)
            cp = _dp(cm)
        self.L.sm_rhs(ctypes.byref(self.g), _dp(w), _dp(bg), t, cp, _dp(r))
        return r

    def run(self, w, t0, dt, nsteps, rk4=True, bg=None):
        bg = np.ascontiguousarray(self.s.bg if bg is None else bg)
        w = self._w(w).copy()
        t = self.L.sm_run(ctypes.byref(self.g), _dp(w), _dp(bg), t0, dt, nsteps, 1 if rk4 else 0)
        return w, t

    @staticmethod
    def threads():
        return lib().sm_num_threads()
