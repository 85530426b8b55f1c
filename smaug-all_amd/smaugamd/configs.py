"""Workload set-ups for the SMAUG per-step MHD update.

Each builder returns a :class:`Setup`: the grid description (the fields of
the reference's ``struct params`` that the stepping path reads, set the way
``iosmaugparams.h`` sets them) and the initial FP64 state in the reference's
slot layout (``wmod``: rho', m_1..m_D, e', b'_1..b'_D, then the background
e_b, rho_b, b_b1..b_bD; cell index i + n0*(j + n1*k), two ghost layers per
side, ``dev/smaug/include/iotypes.h:20-43``).

Workloads:
  brio_wu        1-D Brio-Wu shock tube (carried on the 2-D code path with
                 four interior cells across, as models/iosmaugparams.h.bw does)
  orszag_tang    2-D periodic Orszag-Tang vortex (models/iosmaugparams.h.ot)
  stratified     3-D gravitationally stratified atmosphere, hydrostatic
                 background built in FP64, vertical velocity driver
  alfven_member  one 2-D member of the uneven-size ensemble
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

MU0 = 4.0e-7 * math.pi
RGAS = 8.31e3  # gas constant per kmol with mean molecular weight 1


@dataclasses.dataclass
class Setup:
    name: str
    ndim: int
    n: tuple            # cells per dim incl. ghosts (n[2] == 1 for 2-D)
    bc: tuple           # boundtype per dim: 0 periodic, 4 contcd4
    gamma: float
    g: tuple
    chyp: np.ndarray    # 13 entries, reference field enum order
    hyperdif: int
    driver: int
    drv: np.ndarray     # 8 entries
    dx: list            # per-dim cell metric arrays
    pos: list           # per-dim cell centre arrays
    w: np.ndarray       # evolved fields, shape (NVE, ncell)
    bg: np.ndarray      # background fields, shape (NB, ncell)
    dt: float = 0.0
    nsteps: int = 0
    meta: dict = dataclasses.field(default_factory=dict)

    @property
    def ncell(self) -> int:
        return int(np.prod(self.n))

    @property
    def ninterior(self) -> int:
        m = 1
        for d in range(self.ndim):
            m *= self.n[d] - 4
        return m

    @property
    def nve(self) -> int:
        return 2 + 2 * self.ndim

    @property
    def nbg(self) -> int:
        return 2 + self.ndim

    def field_index(self, name: str) -> int:
        D = self.ndim
        names = ["rho"] + [f"m{d+1}" for d in range(D)] + ["e"] + [f"b{d+1}" for d in range(D)]
        return names.index(name)


def _chyp(ndim, rho=0.02, energy=0.02, b=0.02, mom=0.4):
    """p->chyp[] in the reference enum order (include/iosmaugparams.h:150-172)."""
    c = np.zeros(13)
    D = ndim
    c[0] = rho
    for d in range(D):
        c[1 + d] = mom
        c[2 + D + d] = b
    c[1 + D] = energy
    return c


def _axis(ncell_int, lo, hi):
    """cell centres and metric of a uniform axis with two ghosts per side"""
    h = (hi - lo) / ncell_int
    i = np.arange(ncell_int + 4, dtype=np.float64)
    pos = lo + (i - 2 + 0.5) * h
    dx = np.full(ncell_int + 4, h)
    return pos, dx


def _empty(ndim, n):
    nc = int(np.prod(n))
    return np.zeros((2 + 2 * ndim, nc)), np.zeros((2 + ndim, nc))


def _mesh(pos, n):
    """coordinate arrays broadcast to the (k, j, i) cell grid, flattened i-fastest"""
    n0, n1, n2 = n
    X = np.broadcast_to(pos[0][None, None, :], (n2, n1, n0)).reshape(-1)
    Y = np.broadcast_to(pos[1][None, :, None], (n2, n1, n0)).reshape(-1)
    Z = None
    if len(pos) > 2:
        Z = np.broadcast_to(pos[2][:, None, None], (n2, n1, n0)).reshape(-1)
    return X, Y, Z


def brio_wu(ncells=1024, ny=4):
    """Brio-Wu shock tube on x in [0,1]: left rho=1, p=1, right rho=0.125,
    p=0.1, v=0, bx=0.75, by=+1/-1, gamma=2 (configs/zero1_BW_asc.ini holds the
    same state: e = 1.78125 left, 0.88125 right)."""
    D = 2
    px, dxx = _axis(ncells, 0.0, 1.0)
    py, dxy = _axis(ny, 0.0, ny / ncells)
    n = (ncells + 4, ny + 4, 1)
    w, bg = _empty(D, n)
    X, _, _ = _mesh([px, py], n)
    gam = 2.0
    left = X < 0.5
    rho = np.where(left, 1.0, 0.125)
    p = np.where(left, 1.0, 0.1)
    bx = 0.75
    by = np.where(left, 1.0, -1.0)
    w[0] = rho
    w[3] = p / (gam - 1) + 0.5 * (bx * bx + by * by)
    w[4] = bx
    w[5] = by
    return Setup("brio_wu", D, n, (4, 4), gam, (0.0, 0.0, 0.0), _chyp(D), 1, 0,
                 np.zeros(8), [dxx, dxy], [px, py], w, bg,
                 meta={"t_end": 0.1, "courant": 0.3})


def orszag_tang(nx=1024, ny=None, b0=None, gamma=5.0 / 3.0, chyp_rho=0.045):
    """Orszag-Tang vortex on [0,2pi]^2, periodic: rho=25/(36pi), p=5/(12pi),
    v=(-sin y, sin x), b=b0(-sin y, sin 2x), b0=1/sqrt(4pi); chyp as in
    models/iosmaugparams.h.ot (rho 0.045)."""
    ny = nx if ny is None else ny
    b0 = 1.0 / math.sqrt(4 * math.pi) if b0 is None else b0
    D = 2
    px, dxx = _axis(nx, 0.0, 2 * math.pi)
    py, dxy = _axis(ny, 0.0, 2 * math.pi)
    n = (nx + 4, ny + 4, 1)
    w, bg = _empty(D, n)
    X, Y, _ = _mesh([px, py], n)
    rho = 25.0 / (36.0 * math.pi)
    p = 5.0 / (12.0 * math.pi)
    vx, vy = -np.sin(Y), np.sin(X)
    bx, by = -b0 * np.sin(Y), b0 * np.sin(2 * X)
    w[0] = rho
    w[1] = rho * vx
    w[2] = rho * vy
    w[3] = p / (gamma - 1) + 0.5 * rho * (vx * vx + vy * vy) + 0.5 * (bx * bx + by * by)
    w[4] = bx
    w[5] = by
    return Setup("orszag_tang", D, n, (0, 0), gamma, (0.0, 0.0, 0.0),
                 _chyp(D, rho=chyp_rho), 1, 0, np.zeros(8), [dxx, dxy], [px, py], w, bg,
                 dt=1.0e-4, nsteps=500)


def hydrostatic_profile(z, rho0=2.7e-4, g=274.0, T_lo=6.0e3, T_hi=1.0e6, z_tr=2.0e6,
                        w_tr=0.2e6, sub=64):
    """FP64 hydrostatic equilibrium dp/dz = -rho g, p = rho R T with
    T(z) = T_lo + (T_hi - T_lo)(1 + tanh((z - z_tr)/w_tr))/2 and rho(0) = rho0.
    ln p is integrated with composite Simpson's rule (sub panels per cell)."""
    def temp(zz):
        return T_lo + (T_hi - T_lo) * 0.5 * (1.0 + np.tanh((zz - z_tr) / w_tr))

    z = np.asarray(z, dtype=np.float64)
    order = np.argsort(z)
    zs = z[order]
    # integrate 1/T from 0 to each z
    def integral(a, b):
        m = 2 * sub
        xs = np.linspace(a, b, m + 1)
        f = 1.0 / temp(xs)
        h = (b - a) / m
        return h / 3.0 * (f[0] + f[-1] + 4.0 * f[1:-1:2].sum() + 2.0 * f[2:-1:2].sum())

    I = np.array([integral(0.0, zz) for zz in zs])
    lnp0 = math.log(rho0 * RGAS * temp(np.array(0.0)))
    lnp = lnp0 - g / RGAS * I
    p = np.empty_like(z)
    p[order] = np.exp(lnp)
    T = temp(z)
    rho = p / (RGAS * T)
    return rho, p, T


def stratified(nz=256, nx=256, ny=256, lz=6.0e6, lx=4.0e6, ly=4.0e6, bz_tesla=0.01,
               gamma=5.0 / 3.0, amp=350.0, period=30.0, sigma_h=0.1e6, z_drv=0.1e6,
               sigma_z=0.05e6, courant=0.3):
    """Gravitationally stratified solar atmosphere, vertical axis = dim 0
    (the reference's i axis carries g[0] = -274, include/iosmaugparams.h:112).
    Background rho_b, e_b from the FP64 hydrostatic profile; uniform vertical
    b_b = B/sqrt(mu0) (b^2/2 is the magnetic pressure in code units);
    perturbations start at 0; driver v_z = A sin(2 pi t/P)
    exp(-r^2/sigma_h^2) exp(-(z-z_drv)^2/sigma_z^2) centred horizontally.
    dt is the smaller of courant / max(cfast_d/dx_d) of the initial state and
    the hyperviscous limit 0.25 dx^2 / nu_max (SAC getdt_visc,
    vacusr.viscosity.t:512-546) with nu_max = dx cmax max(chyp) R_max and the
    bound R_max = 6 of the hyperviscosity ratio max|D3|/max|D1|."""
    D = 3
    pz, dz = _axis(nz, 0.0, lz)
    px, dxx = _axis(nx, 0.0, lx)
    py, dxy = _axis(ny, 0.0, ly)
    n = (nz + 4, nx + 4, ny + 4)
    w, bg = _empty(D, n)
    rho_z, p_z, T_z = hydrostatic_profile(pz)
    bb = bz_tesla / math.sqrt(MU0)
    n0, n1, n2 = n
    rho3 = np.broadcast_to(rho_z[None, None, :], (n2, n1, n0)).reshape(-1)
    p3 = np.broadcast_to(p_z[None, None, :], (n2, n1, n0)).reshape(-1)
    bg[0] = p3 / (gamma - 1) + 0.5 * bb * bb   # e_b (thermal + magnetic)
    bg[1] = rho3                                # rho_b
    bg[2] = bb                                  # b_b along z
    drv = np.zeros(8)
    drv[:7] = [amp, period, sigma_h, z_drv, sigma_z, 0.5 * lx, 0.5 * ly]
    s = Setup("stratified", D, n, (4, 4, 4), gamma, (-274.0, 0.0, 0.0), _chyp(D), 1, 1, drv,
              [dz, dxx, dxy], [pz, px, py], w, bg, nsteps=1000,
              meta={"courant": courant, "bz_code": bb})
    s.dt = min(courant / initial_max_courant(s), hyperviscous_dt(s))
    return s


def torsional_driver(s: Setup, amp=500.0, period=30.0, sigma_h=0.1e6, z0=0.1e6, sigma_z=0.05e6,
                     xc=None, yc=None, length=4.0e6) -> Setup:
    """the set-up with the reference's own user driver in place of the
    vertical one: the torsional velocities of dev/smaug/src/usersource.cu:4-86,
    vx = (A yp/L) G sin(2 pi t/P) along dim 1, vy = -(A xp/L) G sin(2 pi t/P)
    along dim 2, G = exp(-(xp^2 + yp^2)/sigma_h^2) exp(-zp^2/sigma_z^2), with
    the reference's constants (A = 500 m/s, P = 30 s, sigma_h = 0.1 Mm,
    z0 = 0.1 Mm, sigma_z = 0.05 Mm, L = 4 Mm) and the centre (xc, yc) in the
    middle of the horizontal extent unless given (the reference's 2 Mm, 2 Mm
    on its 4 x 4 Mm box)"""
    if s.ndim != 3:
        raise ValueError("the torsional driver drives m_y and m_z: 3-D grids only")
    p1, p2 = s.pos[1], s.pos[2]
    if xc is None:
        xc = 0.5 * (p1[2] + p1[-3])
    if yc is None:
        yc = 0.5 * (p2[2] + p2[-3])
    out = dataclasses.replace(s, drv=np.array([amp, period, sigma_h, z0, sigma_z, xc, yc, length]))
    out.driver = 2
    return out


def hyperviscous_dt(s: Setup, rmax: float = 6.0, dtdiffpar: float = 0.25) -> float:
    """0.25 dx^2/nu_max with nu_max = dx cmax_d max(chyp) rmax, over dims,
    for the set-up's initial state (cmax_d from the same fast speed as
    initial_max_courant)"""
    best = float("inf")
    cm = initial_cmax(s)
    ch = float(np.max(s.chyp))
    for d in range(s.ndim):
        dxm = float(np.min(s.dx[d]))
        nu = dxm * cm[d] * ch * rmax
        if nu > 0:
            best = min(best, dtdiffpar * dxm * dxm / nu)
    return best


def initial_cmax(s: Setup):
    """cmax_d over all cells of the set-up's state"""
    return [float(np.max(c)) for c in _cfast(s)]


def _cfast(s: Setup):
    D = s.ndim
    w, bg = s.w, s.bg
    gam = s.gamma
    rho = w[0] + bg[1]
    m = [w[1 + d] for d in range(D)]
    b = [w[2 + D + d] for d in range(D)]
    B = [bg[2 + d] for d in range(D)]
    e = w[1 + D]
    mm = sum(x * x for x in m)
    bb = sum(x * x for x in b)
    bgb = sum(B[d] * b[d] for d in range(D))
    bbb = sum(x * x for x in B)
    pk = (gam - 1) * (e - 0.5 * (mm / rho) - 0.5 * bb - bgb)
    cs2 = (gam * (pk + (gam - 1) * (bg[0] - 0.5 * bbb))) / rho
    cf2 = ((bb + bbb + 2.0 * bgb) / rho) + cs2
    out = []
    for d in range(D):
        btd = B[d] + b[d]
        cfi = 0.5 * (cf2 + np.sqrt(cf2 * cf2 - 4.0 * cs2 * (btd * btd) / rho))
        out.append(np.sqrt(cfi) + np.abs(m[d] / rho))
    return out


def initial_max_courant(s: Setup) -> float:
    """max over dims and cells of cfast_d/dx_d for the set-up's state (FP64, numpy;
    the same expression as computec3, dev/smaug/src/dervfields.cu:418-486)"""
    D = s.ndim
    w, bg = s.w, s.bg
    gam = s.gamma
    rho = w[0] + bg[1]
    m = [w[1 + d] for d in range(D)]
    b = [w[2 + D + d] for d in range(D)]
    B = [bg[2 + d] for d in range(D)]
    e = w[1 + D]
    mm = sum(x * x for x in m)
    bb = sum(x * x for x in b)
    bgb = sum(B[d] * b[d] for d in range(D))
    bbb = sum(x * x for x in B)
    pk = (gam - 1) * (e - 0.5 * (mm / rho) - 0.5 * bb - bgb)
    cs2 = (gam * (pk + (gam - 1) * (bg[0] - 0.5 * bbb))) / rho
    cf2 = ((bb + bbb + 2.0 * bgb) / rho) + cs2
    best = 0.0
    n = s.n
    for d in range(D):
        btd = B[d] + b[d]
        cfi = 0.5 * (cf2 + np.sqrt(cf2 * cf2 - 4.0 * cs2 * (btd * btd) / rho))
        c = np.sqrt(cfi) + np.abs(m[d] / rho)
        shape = (n[2], n[1], n[0])
        dxd = s.dx[d]
        if d == 0:
            dxb = dxd[None, None, :]
        elif d == 1:
            dxb = dxd[None, :, None]
        else:
            dxb = dxd[:, None, None]
        best = max(best, float(np.max(c.reshape(shape) / dxb)))
    return best


def alfven_member(side, b0_scale, gamma=5.0 / 3.0):
    """One member of the uneven-size 2-D ensemble: an Orszag-Tang-type state
    of the given side length with b0 = b0_scale/sqrt(4 pi); dt = 0.3 CFL."""
    s = orszag_tang(nx=side, ny=side, b0=b0_scale / math.sqrt(4 * math.pi), gamma=gamma)
    s.name = f"alfven_{side}"
    s.dt = 0.3 / initial_max_courant(s)
    s.nsteps = 200
    return s


def ensemble_sizes(count=64, seed=0):
    """side lengths drawn uniformly from {96,128,192,256,384,512} and
    b0 scales from U[0.5, 1.5], numpy default_rng(seed)"""
    rng = np.random.default_rng(seed)
    choices = np.array([96, 128, 192, 256, 384, 512])
    sides = choices[rng.integers(0, 6, size=count)]
    scales = rng.uniform(0.5, 1.5, size=count)
    return [int(x) for x in sides], [float(x) for x in scales]


def by_name(name: str, **kw) -> Setup:
    return {"brio_wu": brio_wu, "orszag_tang": orszag_tang, "stratified": stratified}[name](**kw)
