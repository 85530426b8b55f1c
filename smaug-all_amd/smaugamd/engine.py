"""ctypes binding of the MI355X engine (include/smaug_engine.h).

The library is loaded from the package's in-tree build
(``smaug-all_amd/lib/libsmaug_engine.so``).  There is no fallback: if the
library is missing or no HIP device is present the calls raise.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from .configs import Setup

_HERE = os.path.dirname(os.path.abspath(__file__))
PKG_ROOT = os.path.dirname(_HERE)
LIB_PATH = os.path.join(PKG_ROOT, "lib", "libsmaug_engine.so")

FP64 = 0
MIXED = 1

EXPORTS = [
    "smaug_last_error", "smaug_create", "smaug_destroy", "smaug_set_state", "smaug_get_state",
    "smaug_rhs", "smaug_cmax", "smaug_maxcourant", "smaug_rk4", "smaug_euler",
    "smaug_time_rk4", "smaug_ncell", "smaug_ninterior", "smaug_copy_probe", "smaug_stream_probe",
    "smaug_walk_probe",
    "smaug_set_panel", "smaug_get_panel", "smaug_time_pass", "smaug_autotune",
    "smaug_set_zstream", "smaug_get_zstream", "smaug_time_passes", "smaug_set_ratio_form",
    "smaug_get_ratio_form", "smaug_set_bg_plane", "smaug_get_bg_plane",
    "smaug_sums", "smaug_sums_noalloc", "smaug_ens_sums", "smaug_ens_time_passes",
    "smaug_ens_create", "smaug_ens_destroy", "smaug_ens_nrun", "smaug_ens_ninterior",
    "smaug_ens_set_state", "smaug_ens_get_state", "smaug_ens_rk4", "smaug_ens_time_rk4",
    "smaug_ens_set_ratio_form", "smaug_ens_get_ratio_form",
]


class SmaugCfg(ctypes.Structure):
    _fields_ = [
        ("ndim", ctypes.c_int),
        ("n", ctypes.c_int * 3),
        ("bc", ctypes.c_int * 3),
        ("hyperdif", ctypes.c_int),
        ("driver", ctypes.c_int),
        ("precision", ctypes.c_int),
        ("gamma", ctypes.c_double),
        ("g", ctypes.c_double * 3),
        ("chyp", ctypes.c_double * 13),
        ("drv", ctypes.c_double * 8),
        ("dx", ctypes.POINTER(ctypes.c_double) * 3),
        ("pos", ctypes.POINTER(ctypes.c_double) * 3),
    ]


class SmaugError(RuntimeError):
    pass


_lib = None


def load_library(path: str = LIB_PATH):
    global _lib
    if _lib is None:
        if not os.path.exists(path):
            raise SmaugError(f"HIP engine library not built: {path} (run __graft_entry__.build())")
        lib = ctypes.CDLL(path)
        lib.smaug_last_error.restype = ctypes.c_char_p
        lib.smaug_create.argtypes = [ctypes.POINTER(SmaugCfg), ctypes.POINTER(ctypes.c_void_p)]
        lib.smaug_destroy.argtypes = [ctypes.c_void_p]
        dp = ctypes.POINTER(ctypes.c_double)
        lib.smaug_set_state.argtypes = [ctypes.c_void_p, dp, dp]
        lib.smaug_get_state.argtypes = [ctypes.c_void_p, dp]
        lib.smaug_rhs.argtypes = [ctypes.c_void_p, ctypes.c_double, dp]
        lib.smaug_cmax.argtypes = [ctypes.c_void_p, dp]
        lib.smaug_maxcourant.argtypes = [ctypes.c_void_p, dp]
        lib.smaug_rk4.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_double, ctypes.c_int]
        lib.smaug_euler.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_double, ctypes.c_int]
        lib.smaug_time_rk4.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_double,
                                       ctypes.c_int, dp, dp]
        lib.smaug_ncell.argtypes = [ctypes.c_void_p]
        lib.smaug_ncell.restype = ctypes.c_longlong
        lib.smaug_ninterior.argtypes = [ctypes.c_void_p]
        lib.smaug_ninterior.restype = ctypes.c_longlong
        lib.smaug_copy_probe.argtypes = [ctypes.c_longlong, ctypes.c_int, dp]
        lib.smaug_stream_probe.argtypes = [ctypes.c_longlong, ctypes.c_int, dp]
        lib.smaug_walk_probe.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                         ctypes.c_int, dp]
        vp = ctypes.c_void_p
        ip = ctypes.POINTER(ctypes.c_int)
        lib.smaug_set_panel.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        lib.smaug_get_panel.argtypes = [vp, ctypes.c_int, ip, ip]
        lib.smaug_time_pass.argtypes = [vp, ctypes.c_int, ctypes.c_int, dp]
        lib.smaug_autotune.argtypes = [vp, ctypes.c_int, ip, dp]
        lib.smaug_set_zstream.argtypes = [vp, ctypes.c_int, ctypes.c_int]
        lib.smaug_get_zstream.argtypes = [vp, ip, ip]
        lib.smaug_set_ratio_form.argtypes = [vp, ctypes.c_int, ctypes.c_int]
        lib.smaug_get_ratio_form.argtypes = [vp, ip, ip]
        lib.smaug_set_bg_plane.argtypes = [vp, ctypes.c_int]
        lib.smaug_get_bg_plane.argtypes = [vp, ip]
        lib.smaug_sums.argtypes = [vp, dp]
        lib.smaug_sums_noalloc.argtypes = [ctypes.c_int, ip, ctypes.c_int, vp, vp, vp]
        lib.smaug_ens_sums.argtypes = [vp, ctypes.c_int, dp]
        lib.smaug_ens_time_passes.argtypes = [vp, dp, dp, dp]
        lib.smaug_time_passes.argtypes = [vp, ctypes.c_double, ctypes.c_double, dp]
        lib.smaug_ens_create.argtypes = [ctypes.POINTER(SmaugCfg), ctypes.c_int,
                                         ctypes.POINTER(vp)]
        lib.smaug_ens_destroy.argtypes = [vp]
        lib.smaug_ens_nrun.argtypes = [vp]
        lib.smaug_ens_ninterior.argtypes = [vp]
        lib.smaug_ens_ninterior.restype = ctypes.c_longlong
        lib.smaug_ens_set_state.argtypes = [vp, ctypes.c_int, dp, dp]
        lib.smaug_ens_get_state.argtypes = [vp, ctypes.c_int, dp]
        lib.smaug_ens_rk4.argtypes = [vp, dp, dp, ctypes.c_int]
        lib.smaug_ens_time_rk4.argtypes = [vp, dp, dp, ctypes.c_int, dp]
        lib.smaug_ens_set_ratio_form.argtypes = [vp, ctypes.c_int]
        lib.smaug_ens_get_ratio_form.argtypes = [vp, ip]
        _lib = lib
    return _lib


def sums_noalloc(ndim: int, n, precision: int, d_w: int, d_sums: int, stream: int = 0) -> int:
    """smaug_sums_noalloc on device addresses (integers, e.g. a torch
    tensor's data_ptr()); returns the library's status code"""
    lib = load_library()
    nn = (ctypes.c_int * 3)(*[int(x) for x in (list(n) + [1, 1, 1])[:3]])
    return lib.smaug_sums_noalloc(ndim, nn, precision, ctypes.c_void_p(d_w),
                                  ctypes.c_void_p(d_sums), ctypes.c_void_p(stream or None))


def last_error() -> str:
    return load_library().smaug_last_error().decode()


def _dptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def make_cfg(s: Setup, precision: int):
    """the smaug_cfg for a set-up; returns (cfg, keepalive arrays)"""
    cfg = SmaugCfg()
    cfg.ndim = s.ndim
    keep = []
    for d in range(3):
        cfg.n[d] = s.n[d] if d < len(s.n) else 1
        cfg.bc[d] = s.bc[d] if d < s.ndim else 0
        cfg.g[d] = s.g[d]
    cfg.hyperdif = s.hyperdif
    cfg.driver = s.driver
    cfg.precision = precision
    cfg.gamma = s.gamma
    for f in range(13):
        cfg.chyp[f] = float(s.chyp[f])
    for k in range(8):
        cfg.drv[k] = float(s.drv[k])
    for d in range(s.ndim):
        dx = np.ascontiguousarray(s.dx[d], dtype=np.float64)
        pos = np.ascontiguousarray(s.pos[d], dtype=np.float64)
        keep += [dx, pos]
        cfg.dx[d] = _dptr(dx)
        cfg.pos[d] = _dptr(pos)
    return cfg, keep


class Engine:
    """One grid on the GPU.  All state transfer is FP64 on the host side."""

    def __init__(self, setup: Setup, precision: int = FP64):
        self.lib = load_library()
        self.setup = setup
        self.precision = precision
        cfg, self._keep = make_cfg(setup, precision)
        h = ctypes.c_void_p()
        self._check(self.lib.smaug_create(ctypes.byref(cfg), ctypes.byref(h)))
        self.h = h

    def _check(self, rc):
        if rc != 0:
            raise SmaugError(f"smaug error {rc}: {self.lib.smaug_last_error().decode()}")

    def close(self):
        if getattr(self, "h", None):
            self.lib.smaug_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_state(self, w=None, bg=None):
        s = self.setup
        w = np.ascontiguousarray(s.w if w is None else w, dtype=np.float64)
        bg = np.ascontiguousarray(s.bg if bg is None else bg, dtype=np.float64)
        assert w.shape == (s.nve, s.ncell) and bg.shape == (s.nbg, s.ncell)
        self._check(self.lib.smaug_set_state(self.h, _dptr(w), _dptr(bg)))

    def get_state(self):
        s = self.setup
        w = np.empty((s.nve, s.ncell))
        self._check(self.lib.smaug_get_state(self.h, _dptr(w)))
        return w

    def rhs(self, t=0.0):
        s = self.setup
        r = np.empty((s.nve, s.ncell))
        self._check(self.lib.smaug_rhs(self.h, t, _dptr(r)))
        return r

    def cmax(self):
        c = np.zeros(3)
        self._check(self.lib.smaug_cmax(self.h, _dptr(c)))
        return c

    def maxcourant(self):
        """max over cells and dims of cfast_d / dx_d of the current state"""
        v = np.zeros(1)
        self._check(self.lib.smaug_maxcourant(self.h, _dptr(v)))
        return float(v[0])

    def rk4(self, t0, dt, nsteps):
        self._check(self.lib.smaug_rk4(self.h, t0, dt, nsteps))

    def euler(self, t0, dt, nsteps):
        self._check(self.lib.smaug_euler(self.h, t0, dt, nsteps))

    def time_rk4(self, t0, dt, nsteps):
        tot = ctypes.c_double()
        st = ctypes.c_double()
        self._check(self.lib.smaug_time_rk4(self.h, t0, dt, nsteps, ctypes.byref(tot),
                                            ctypes.byref(st)))
        return tot.value, st.value


    # launch-shape tuning: pass 0 = hyperviscosity ratios, 1 = stage update
    def set_panel(self, pass_, tpx, tpy):
        self._check(self.lib.smaug_set_panel(self.h, pass_, tpx, tpy))

    def get_panel(self, pass_):
        x, y = ctypes.c_int(), ctypes.c_int()
        self._check(self.lib.smaug_get_panel(self.h, pass_, ctypes.byref(x), ctypes.byref(y)))
        return x.value, y.value

    def set_zstream(self, enable, kz=0):
        """3-D: z-streamed stage update (default) or the per-cell one"""
        self._check(self.lib.smaug_set_zstream(self.h, int(enable), kz))

    def sums(self):
        """compensated FP64 sums of the evolved fields over the interior"""
        out = np.zeros(self.setup.nve)
        self._check(self.lib.smaug_sums(self.h, _dptr(out)))
        return out

    PASSES = ("aux+cmax", "ratio", "hyperdiffusive sources", "stage update", "ghost fill")

    def time_passes(self, t0, dt):
        """one more RK4 step; mean ms per launch of each pass (PASSES order)"""
        ms = (ctypes.c_double * 5)()
        self._check(self.lib.smaug_time_passes(self.h, t0, dt, ms))
        return dict(zip(self.PASSES, list(ms)))

    def set_ratio_form(self, zstreamed, kz=0):
        """3-D: hyperviscosity ratios z-streamed (1), per cell (0), or
        z-streamed along k with the in-plane pairs formed by the z-streamed
        source kernel (2, the default).  2-D: every pair formed by the update
        kernel (2, the default) or by the per-cell ratio pass (0)"""
        self._check(self.lib.smaug_set_ratio_form(self.h, int(zstreamed), kz))

    def get_ratio_form(self):
        """(form of the 3-D ratio pass: 0 per cell, 1 z-streamed, 2 z-streamed
        pairs along k with the in-plane pairs formed by the source kernel; its
        z-stream depth)"""
        zs, kz = ctypes.c_int(), ctypes.c_int()
        self._check(self.lib.smaug_get_ratio_form(self.h, ctypes.byref(zs), ctypes.byref(kz)))
        return zs.value, kz.value

    def get_zstream(self):
        """(form of the 3-D update pass: 0 per cell, 1 z-streamed, 2 z-streamed
        with per-cell hyperdiffusion; z-stream depth)"""
        en, kz = ctypes.c_int(), ctypes.c_int()
        self._check(self.lib.smaug_get_zstream(self.h, ctypes.byref(en), ctypes.byref(kz)))
        return en.value, kz.value

    def set_bg_plane(self, enable):
        """3-D: read a background that is the same on every k-plane from its
        one-plane copy (default) or always from the full arrays"""
        self._check(self.lib.smaug_set_bg_plane(self.h, int(enable)))

    def get_bg_plane(self):
        """True if the kernels read the background from its k-plane copy"""
        a = ctypes.c_int()
        self._check(self.lib.smaug_get_bg_plane(self.h, ctypes.byref(a)))
        return bool(a.value)

    def time_pass(self, pass_, reps=5):
        ms = ctypes.c_double()
        self._check(self.lib.smaug_time_pass(self.h, pass_, reps, ctypes.byref(ms)))
        return ms.value

    def autotune(self, reps=3):
        """fastest XCD panel per pass; returns ({pass: (tpx, tpy)}, {pass: ms})"""
        best = (ctypes.c_int * 4)()
        ms = (ctypes.c_double * 2)()
        self._check(self.lib.smaug_autotune(self.h, reps, best, ms))
        return ({0: (best[0], best[1]), 1: (best[2], best[3])}, {0: ms[0], 1: ms[1]})


class Ensemble:
    """Uneven grids advanced together by the persistent queue-driven kernels.
    Every set-up keeps its own grid, time and dt; the physics parameters
    (ndim, hyperdif, driver, gamma, g, chyp, drv) must agree."""

    def __init__(self, setups, precision: int = MIXED):
        self.lib = load_library()
        self.setups = list(setups)
        self.precision = precision
        arr = (SmaugCfg * len(self.setups))()
        self._keep = []
        for r, s in enumerate(self.setups):
            cfg, keep = make_cfg(s, precision)
            arr[r] = cfg
            self._keep += keep
        self._cfgs = arr
        h = ctypes.c_void_p()
        self._check(self.lib.smaug_ens_create(arr, len(self.setups), ctypes.byref(h)))
        self.h = h

    def _check(self, rc):
        if rc != 0:
            raise SmaugError(f"smaug error {rc}: {self.lib.smaug_last_error().decode()}")

    def close(self):
        if getattr(self, "h", None):
            self.lib.smaug_ens_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def ninterior(self) -> int:
        return int(self.lib.smaug_ens_ninterior(self.h))

    def set_state(self, run, w=None, bg=None):
        s = self.setups[run]
        w = np.ascontiguousarray(s.w if w is None else w, dtype=np.float64)
        bg = np.ascontiguousarray(s.bg if bg is None else bg, dtype=np.float64)
        assert w.shape == (s.nve, s.ncell) and bg.shape == (s.nbg, s.ncell)
        self._check(self.lib.smaug_ens_set_state(self.h, run, _dptr(w), _dptr(bg)))

    def get_state(self, run):
        s = self.setups[run]
        w = np.empty((s.nve, s.ncell))
        self._check(self.lib.smaug_ens_get_state(self.h, run, _dptr(w)))
        return w

    def sums(self, run):
        """compensated FP64 sums of run's evolved fields over its interior"""
        out = np.zeros(self.setups[run].nve)
        self._check(self.lib.smaug_ens_sums(self.h, run, _dptr(out)))
        return out

    def _times(self, t0, dt):
        n = len(self.setups)
        t0 = np.ascontiguousarray(np.broadcast_to(np.asarray(t0, dtype=np.float64), (n,)))
        dt = np.ascontiguousarray(np.broadcast_to(np.asarray(dt, dtype=np.float64), (n,)))
        return t0, dt

    def rk4(self, t0, dt, nsteps):
        t0, dt = self._times(t0, dt)
        self._check(self.lib.smaug_ens_rk4(self.h, _dptr(t0), _dptr(dt), nsteps))

    def time_passes(self, t0, dt):
        """one more RK4 step; mean ms per launch of each pass"""
        t0, dt = self._times(t0, dt)
        ms = (ctypes.c_double * 4)()
        self._check(self.lib.smaug_ens_time_passes(self.h, _dptr(t0), _dptr(dt), ms))
        return dict(zip(("aux+cmax", "ratio", "stage update", "ghost fill"), list(ms)))

    def time_rk4(self, t0, dt, nsteps):
        t0, dt = self._times(t0, dt)
        ms = ctypes.c_double()
        self._check(self.lib.smaug_ens_time_rk4(self.h, _dptr(t0), _dptr(dt), nsteps,
                                                ctypes.byref(ms)))
        return ms.value

    def set_ratio_form(self, form):
        """2-D runs: the stage update forms the ratios itself (2, the default)
        or reads them from the ratio pass (0)"""
        self._check(self.lib.smaug_ens_set_ratio_form(self.h, int(form)))

    def get_ratio_form(self):
        f = ctypes.c_int()
        self._check(self.lib.smaug_ens_get_ratio_form(self.h, ctypes.byref(f)))
        return f.value


def stream_probe(ncell=1 << 24, reps=10):
    """measured bandwidth (GB/s, read + write) of 32 read and 20 write FP32
    streams, one value of each per cell: a mixed 3-D stage update's streams"""
    lib = load_library()
    g = ctypes.c_double()
    rc = lib.smaug_stream_probe(ncell, reps, ctypes.byref(g))
    if rc:
        raise SmaugError(lib.smaug_last_error().decode())
    return g.value


def walk_probe(n, kz=128, reps=10):
    """measured bandwidth (GB/s, read + write of the interior) of the same
    streams walked as the z-streamed stage update walks them: 64 x 4-column
    tiles of a grid of n = (n0, n1, n2) cells, chunks of kz planes"""
    lib = load_library()
    g = ctypes.c_double()
    rc = lib.smaug_walk_probe(int(n[0]), int(n[1]), int(n[2]), int(kz), reps, ctypes.byref(g))
    if rc:
        raise SmaugError(lib.smaug_last_error().decode())
    return g.value


def copy_probe(nbytes=2 << 30, reps=10):
    """measured streaming-copy bandwidth (GB/s, read + write) of the device"""
    lib = load_library()
    g = ctypes.c_double()
    rc = lib.smaug_copy_probe(nbytes, reps, ctypes.byref(g))
    if rc:
        raise SmaugError(lib.smaug_last_error().decode())
    return g.value
