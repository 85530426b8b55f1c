"""Pins of the CPU oracle (oracle/smaug_oracle.c) against closed forms,
invariants, convergence orders and the reference's own Brio-Wu fixture.

Each check is chosen so that a plausible restatement mistake (a dropped
term, a wrong sign, a shifted window or stencil weight) makes it fail.
"""
import json
import math
import os

import numpy as np
import pytest

from helpers import interior_mask, perturbed
from oracle.oracle import Oracle
from smaugamd import configs

HERE = os.path.dirname(os.path.abspath(__file__))


def grid2d(nx=16, ny=12, bc=(0, 0), lx=1.0, ly=1.0, hyperdif=0, gamma=5.0 / 3.0):
    s = configs.orszag_tang(nx=nx, ny=ny)
    px, dxx = configs._axis(nx, 0.0, lx)
    py, dxy = configs._axis(ny, 0.0, ly)
    s.dx, s.pos = [dxx, dxy], [px, py]
    s.bc = bc
    s.hyperdif = hyperdif
    s.gamma = gamma
    return s


def cube3d(n=(12, 10, 11), bc=(0, 0, 0), hyperdif=0):
    s = configs.stratified(nz=n[0], nx=n[1], ny=n[2], lz=1.0, lx=1.0, ly=1.0)
    s.bg[:] = 0.0
    s.driver = 0
    s.g = (0.0, 0.0, 0.0)
    s.bc = bc
    s.hyperdif = hyperdif
    dxs, poss = [], []
    for d in range(3):
        p, h = configs._axis(n[d], 0.0, 1.0)
        dxs.append(h)
        poss.append(p)
    s.dx, s.pos = dxs, poss
    return s


def cell_coords(s):
    n = s.n
    k, j, i = np.meshgrid(np.arange(n[2]), np.arange(n[1]), np.arange(n[0]), indexing="ij")
    return i.reshape(-1), j.reshape(-1), k.reshape(-1)


# ---------------------------------------------------------------- grad4
@pytest.mark.parametrize("d", [0, 1])
def test_grad4_exact_for_quartics(d):
    s = grid2d(nx=20, ny=18, bc=(0, 0))
    o = Oracle(s)
    x = s.pos[d][cell_coords(s)[d]]
    a = [0.3, -1.2, 0.7, 2.0, -0.9]
    q = sum(c * x ** k for k, c in enumerate(a))
    dq = sum(k * c * x ** (k - 1) for k, c in enumerate(a) if k > 0)
    g = o.grad4(q, d)
    xi = cell_coords(s)[d]
    sel = (xi > 1) & (xi < s.n[d] - 2)
    np.testing.assert_allclose(g[sel], dq[sel], rtol=1e-9, atol=1e-9)
    assert np.all(g[~sel] == 0)


def test_grad4_fourth_order_and_edge_rule():
    errs = []
    for nx in (24, 48, 96):
        s = grid2d(nx=nx, ny=8, bc=(4, 4), lx=2 * math.pi)
        o = Oracle(s)
        i, j, _ = cell_coords(s)
        x = s.pos[0][i]
        g = o.grad4(np.sin(x), 0)
        # non-periodic edge rule of grad3dn: x = 2, n-3 always; x = 3, n-4
        # where the transverse index is interior
        jin = (j > 1) & (j < s.n[1] - 2)
        zero = np.isin(i, [2, nx + 4 - 3]) | (np.isin(i, [3, nx + 4 - 4]) & jin)
        assert np.all(g[zero] == 0)
        assert np.all(g[np.isin(i, [3, nx]) & ~jin] != 0)
        sel = (i > 3) & (i < nx + 4 - 4)
        errs.append(np.abs(g[sel] - np.cos(x[sel])).max())
    r1, r2 = errs[0] / errs[1], errs[1] / errs[2]
    assert 13 < r1 < 19 and 13 < r2 < 19, errs


# ---------------------------------------------------------------- boundary
def test_boundary_periodic_and_mirror_images():
    s = grid2d(nx=16, ny=12, bc=(0, 4), lx=1.0, ly=1.0)
    o = Oracle(s)
    i, j, _ = cell_coords(s)
    x, y = s.pos[0][i], s.pos[1][j]
    yc = s.pos[1][2]  # contcd4 mirrors about the first interior cell
    yh = s.pos[1][s.n[1] - 3]  # and the last interior cell
    f = np.sin(2 * math.pi * x) * np.where(y < 0.5, np.cos(7 * (y - yc)), np.cos(5 * (y - yh)))
    w = np.zeros((s.nve, s.ncell))
    w[0] = f
    m = interior_mask(s)
    w[0][~m] = 1e9
    got = o.boundary(w)[0]
    np.testing.assert_allclose(got, f, rtol=0, atol=1e-12)


# ---------------------------------------------------------------- cfast / cmax
def test_cmax_matches_textbook_fast_speed_under_any_split():
    rng = np.random.default_rng(5)
    gam = 5.0 / 3.0
    rho, p = 1.7, 0.9
    v = np.array([0.3, -0.2])
    B = np.array([0.8, 0.5])
    cs2 = gam * p / rho
    va2 = B @ B / rho
    for d in range(2):
        cf2 = 0.5 * (cs2 + va2 + math.sqrt((cs2 + va2) ** 2 - 4 * cs2 * B[d] ** 2 / rho))
        expect = math.sqrt(cf2) + abs(v[d])
        for trial in range(3):
            s = grid2d(nx=8, ny=8, gamma=gam)
            rb = rng.uniform(0, rho)
            Bb = rng.uniform(-1, 1, 2)
            pb = rng.uniform(0, p)
            eb = pb / (gam - 1) + 0.5 * Bb @ Bb
            etot = p / (gam - 1) + 0.5 * rho * v @ v + 0.5 * B @ B
            s.bg[0], s.bg[1], s.bg[2], s.bg[3] = eb, rb, Bb[0], Bb[1]
            s.w[0] = rho - rb
            s.w[1], s.w[2] = rho * v[0], rho * v[1]
            s.w[3] = etot - eb
            s.w[4], s.w[5] = B[0] - Bb[0], B[1] - Bb[1]
            c = Oracle(s).cmax(s.w)
            assert abs(c[d] - expect) < 1e-12 * expect


def test_cmax_includes_ghost_cells():
    s = grid2d(nx=8, ny=8)
    o = Oracle(s)
    w = o.boundary(s.w)
    base = o.cmax(w)
    w2 = w.copy()
    w2[1][0] = 50.0 * w2[0][0]  # cell (0,0) is a ghost corner
    assert o.cmax(w2)[0] > base[0] + 40


# ---------------------------------------------------------------- hyperviscosity
def _nu_line(q_of_x, n=24, dim=0, cmax=2.0, chyp_rho=0.5):
    s = grid2d(nx=n - 4, ny=8, bc=(4, 4))
    s.chyp[0] = chyp_rho
    s.dx = [np.ones(s.n[0]), np.ones(s.n[1])]
    i, j, _ = cell_coords(s)
    w = np.zeros((s.nve, s.ncell))
    w[0] = q_of_x(i.astype(float))
    w[3] = 10.0
    s.bg[1] = 1.0
    nr, nl = Oracle(s).nu(w, 0, dim, cmax)
    return s, i, j, nr, nl


def test_nu_vanishes_for_quadratics_except_at_the_pads():
    s, i, j, nr, nl = _nu_line(lambda x: 0.3 * x * x - x + 2)
    n = s.n[0]
    inner = (j >= 2) & (j <= s.n[1] - 3)
    interior = inner & (i >= 2) & (i <= n - 3)
    K = 1.0 * 2.0 * 0.5
    quiet = interior & (i > 2) & (i < n - 3)
    # D3 of a quadratic vanishes up to rounding of the differences
    assert np.abs(nr[quiet]).max() < 1e-12 and np.abs(nl[quiet]).max() < 1e-12
    assert np.all(nr[interior & (i == n - 3)] > 0.1) and np.all(nl[interior & (i == 2)] > 0.1)
    assert np.all(nr[~interior] == 0) and np.all(nl[~interior] == 0)


def test_nu_linear_profile_pad_values():
    # q = x: D3 = 0 everywhere except where a window reaches a pad; the pads
    # hold q[5] (left) and q[n-6] (right): nu = K*6/5 at x = 2 (L), n-3 (R)
    s, i, j, nr, nl = _nu_line(lambda x: x)
    n = s.n[0]
    K = 1.0 * 2.0 * 0.5
    row = (j == 4)
    got_r = nr[row & (i == n - 3)][0]
    got_l = nl[row & (i == 2)][0]
    assert abs(got_r - K * 6 / 5) < 1e-14 and abs(got_l - K * 6 / 5) < 1e-14
    others = row & (i > 2) & (i < n - 3)
    assert np.all(nr[others] == 0) and np.all(nl[others] == 0)


def test_nu_cubic_closed_form_window_offsets():
    # q = x^3: D3(f) = 6 for every f, D1(f) = 3f^2+3f+1, so away from the pads
    # nu_R(x) = 6K / D1(x+2) and nu_L(x) = 6K / D1(x+1)
    s, i, j, nr, nl = _nu_line(lambda x: x ** 3, n=28)
    n = s.n[0]
    K = 1.0
    D1 = lambda f: 3 * f * f + 3 * f + 1
    for x in range(3, n - 4):
        r = nr[(j == 4) & (i == x)][0]
        l = nl[(j == 4) & (i == x)][0]
        assert abs(r - 6 * K / D1(x + 2)) < 1e-14 * max(1, r)
        assert abs(l - 6 * K / D1(x + 1)) < 1e-14 * max(1, l)


# ---------------------------------------------------------------- operator
def test_uniform_state_is_steady():
    s = grid2d(nx=12, ny=10, hyperdif=1)
    w = np.zeros((s.nve, s.ncell))
    w[0], w[1], w[2], w[3], w[4], w[5] = 1.3, 0.4, -0.2, 3.0, 0.5, 0.7
    r = Oracle(s).rhs(w)
    assert np.abs(r).max() < 1e-13


def test_periodic_advection_conserves_every_field():
    s = grid2d(nx=24, ny=20, bc=(0, 0))
    s.w = perturbed(s, amp=0.05, seed=3)
    o = Oracle(s)
    w = o.boundary(s.w)
    r = o.rhs(w)
    m = interior_mask(s)
    tot = r[:, m].sum(axis=1)
    scale = np.abs(r[:, m]).sum(axis=1)
    assert np.all(np.abs(tot) <= 1e-13 * scale), (tot, scale)


@pytest.mark.parametrize("hyperdif", [0, 1])
def test_split_invariance_with_uniform_background(hyperdif):
    """The perturbation equations with a uniform background are the full MHD
    equations: moving part of a state into the background leaves d/dt of the
    evolved fields unchanged (energy excepted under hyperdiffusion, whose q is
    e' - b'^2/2 - m^2/2rho)."""
    s = cube3d(n=(10, 9, 8), bc=(0, 0, 0), hyperdif=hyperdif)
    rng = np.random.default_rng(11)
    i, j, k = cell_coords(s)
    x, y, z = s.pos[0][i], s.pos[1][j], s.pos[2][k]
    tp = 2 * math.pi
    rho = 1.0 + 0.2 * np.sin(tp * x) * np.cos(tp * y)
    v = [0.3 * np.sin(tp * y), 0.2 * np.cos(tp * z), 0.1 * np.sin(tp * x)]
    B = [0.5 + 0.1 * np.cos(tp * z), 0.3 * np.sin(tp * x), -0.2 + 0.1 * np.sin(tp * y)]
    p = 1.0 + 0.1 * np.cos(tp * x)
    gam = s.gamma
    etot = p / (gam - 1) + 0.5 * rho * sum(a * a for a in v) + 0.5 * sum(b * b for b in B)
    full = np.zeros((s.nve, s.ncell))
    full[0] = rho
    for d in range(3):
        full[1 + d] = rho * v[d]
        full[5 + d] = B[d]
    full[4] = etot
    s.bg[:] = 0.0
    oa = Oracle(s)
    ra = oa.rhs(oa.boundary(full))
    rb_, Bb, pb = 0.6, rng.uniform(-0.4, 0.4, 3), 0.5
    bg = np.zeros_like(s.bg)
    bg[0] = pb / (gam - 1) + 0.5 * Bb @ Bb
    bg[1] = rb_
    bg[2:5] = Bb[:, None]
    pert = full.copy()
    pert[0] -= rb_
    pert[4] -= bg[0]
    for d in range(3):
        pert[5 + d] -= Bb[d]
    s.bg = bg
    ob = Oracle(s)
    rb = ob.rhs(ob.boundary(pert))
    m = interior_mask(s)
    fields = range(s.nve) if not hyperdif else [0, 1, 2, 3, 5, 6, 7]
    for f in fields:
        scale = np.abs(ra[f][m]).max()
        assert np.abs(ra[f][m] - rb[f][m]).max() < 1e-11 * scale, f


def test_operator_fourth_order_convergence():
    """hyperdiffusion off: L(w) of a smooth periodic state converges at 4th
    order (cell-centred grids refined by 3 share cell centres)"""
    vals = {}
    for nx in (24, 72, 216):
        s = configs.orszag_tang(nx=nx, ny=nx)
        s.hyperdif = 0
        o = Oracle(s)
        r = o.rhs(o.boundary(s.w))
        vals[nx] = r.reshape(6, nx + 4, nx + 4)[:, 2:-2, 2:-2]
    def coarse(a, f):
        return a[:, f // 2::f, f // 2::f]
    e1 = np.abs(vals[24] - coarse(vals[72], 3)).max(axis=(1, 2))
    e2 = np.abs(vals[72] - coarse(vals[216], 3)).max(axis=(1, 2))
    for f in range(1, 6):
        assert 65 < e1[f] / e2[f] < 95, (f, e1, e2)  # 3^4 = 81


def test_gravity_and_driver_closed_forms():
    s = configs.stratified(nz=12, nx=10, ny=10, lz=1.2e6, lx=1e6, ly=1e6)
    s.hyperdif = 0
    o = Oracle(s)
    D = 3
    w = np.zeros((s.nve, s.ncell))
    rng = np.random.default_rng(2)
    w[0] = 1e-6 * rng.uniform(0.5, 1.0, s.ncell) * 0  # no density perturbation first
    t = 7.5
    r = o.rhs(o.boundary(w), t=t)
    i, j, k = cell_coords(s)
    z, x, y = s.pos[0][i], s.pos[1][j], s.pos[2][k]
    A, P, sh, z0, sz, xc, yc = s.drv[:7]
    v = A * math.sin(2 * math.pi * t / P) * np.exp(-(z - z0) ** 2 / sz ** 2) * \
        np.exp(-((x - xc) ** 2 + (y - yc) ** 2) / sh ** 2)
    rho = s.bg[1]
    m = interior_mask(s)
    np.testing.assert_allclose(r[1][m], (v * rho)[m], rtol=1e-12, atol=1e-300)
    np.testing.assert_allclose(r[4][m], (0.5 * v * v * rho)[m], rtol=1e-12, atol=1e-300)
    # gravity: a uniform density perturbation at rest accelerates with g
    s.driver = 0
    o = Oracle(s)
    w[0] = 1e-3 * s.bg[1]
    r = o.rhs(o.boundary(w), t=t)
    np.testing.assert_allclose(r[1][m], (-274.0 * w[0])[m], rtol=1e-12)


def test_torsional_driver_closed_forms():
    # the reference's own driver (usersource.cu:4-86) on a state at rest
    # with no perturbation: L(w) is the driver alone
    s = configs.torsional_driver(configs.stratified(nz=12, nx=12, ny=12, lz=1.2e6, lx=1e6,
                                                    ly=1e6), sigma_h=0.2e6, sigma_z=0.3e6,
                                 z0=0.5e6)
    s.hyperdif = 0
    o = Oracle(s)
    t = 4.0
    r = o.rhs(o.boundary(np.zeros((s.nve, s.ncell))), t=t)
    m = interior_mask(s)
    i, j, k = cell_coords(s)
    A, P, sh, z0, sz, xc, yc, L = s.drv
    xp, yp, zp = s.pos[1][j] - xc, s.pos[2][k] - yc, s.pos[0][i] - z0
    rho = s.bg[1]
    vx, vy = r[2] / rho, r[3] / rho
    scale = np.abs(vx[m]).max()
    assert scale > 0
    # purely horizontal and torsional: no vertical push, no radial component
    assert np.all(r[1][m] == 0)
    np.testing.assert_allclose((vx * xp + vy * yp)[m], 0, atol=1e-12 * scale * 1e6)
    # energy = rho |v|^2 / 2 of the same velocities
    np.testing.assert_allclose(r[4][m], (0.5 * rho * (vx * vx + vy * vy))[m], rtol=1e-12)
    # the grid is symmetric about the centre: no net momentum is injected
    assert abs(r[2][m].sum()) < 1e-12 * np.abs(r[2][m]).sum()
    assert abs(r[3][m].sum()) < 1e-12 * np.abs(r[3][m]).sum()
    # the speed |v| = A r/L exp(-r^2/sh^2) exp(-zp^2/sz^2) |sin(2 pi t/P)|
    rad = np.sqrt(xp * xp + yp * yp)
    speed = A * rad / L * np.exp(-rad ** 2 / sh ** 2) * np.exp(-zp ** 2 / sz ** 2) * \
        abs(math.sin(2 * math.pi * t / P))
    np.testing.assert_allclose(np.hypot(vx, vy)[m], speed[m], rtol=1e-12, atol=1e-300)
    # the rotation sense: vx follows +yp and vy follows -xp (times sin(2 pi t/P))
    td = math.sin(2 * math.pi * t / P)
    live = m & (np.hypot(vx, vy) > 1e-6 * scale)
    assert np.all(np.sign(vx[live] * td) == np.sign(yp[live]))
    assert np.all(np.sign(vy[live] * td) == -np.sign(xp[live]))


def test_rk4_is_fourth_order_in_time():
    s = configs.orszag_tang(nx=16)
    s.hyperdif = 0
    o = Oracle(s)
    w0 = o.boundary(s.w)
    T = 0.08
    ref, _ = o.run(w0, 0.0, T / 64, 64)
    errs = []
    for nst in (4, 8):
        w, _ = o.run(w0, 0.0, T / nst, nst)
        errs.append(np.abs(w - ref).max())
    assert 12 < errs[0] / errs[1] < 20, errs


def test_euler_is_first_order_in_time():
    s = configs.orszag_tang(nx=16)
    s.hyperdif = 0
    o = Oracle(s)
    w0 = o.boundary(s.w)
    T = 0.02
    ref, _ = o.run(w0, 0.0, T / 256, 256, rk4=False)
    errs = []
    for nst in (8, 16):
        w, _ = o.run(w0, 0.0, T / nst, nst, rk4=False)
        errs.append(np.abs(w - ref).max())
    assert 1.6 < errs[0] / errs[1] < 2.4, errs


# ---------------------------------------------------------------- reference fixture
def test_brio_wu_states_match_reference_ini():
    fx = json.load(open(os.path.join(HERE, "golden", "bw_ini_summary.json")))
    assert fx["eqpar"][0] == 2.0  # gamma
    s = configs.brio_wu(ncells=64)
    left, right = np.array(fx["rows"]["left"]), np.array(fx["rows"]["right"])
    # columns h m1 m2 e b1 b2 eb rhob bg1 bg2
    i, j, _ = cell_coords(s)
    cl = (i == 2) & (j == 2)
    cr = (i == s.n[0] - 3) & (j == 2)
    for cell, row in ((cl, left), (cr, right)):
        got = np.concatenate([s.w[:, cell][:, 0], s.bg[:, cell][:, 0]])
        assert np.allclose(got, row, rtol=0, atol=1e-15), (got, row)
    assert s.gamma == fx["eqpar"][0]
    assert fx["y_uniform"] and fx["distinct_rows"] == 2


# ---------------------------------------------------------------- hyperdiffusive sources
# Closed forms on polynomial states of a 3-D box with uniform but unequal
# spacings.  With q = a*x^3 along d (x = cell index) the differences in
# index units are D1(f) = q[f+1]-q[f] = a(3f^2+3f+1) and D3(f) = 6a, so
# away from the pads nu_R(x) = h c chyp 6/D1(x+2) and
# nu_L(x) = h c chyp 6/D1(x+1).  Each source is then written from the
# reference's definition of the term:
#   rho, e:       (nu_R dR q - nu_L dL q)/h       hyperdifrhosource1.cu:15-110,
#                                                  hyperdifesource1.cu:18-545
#   m diagonal:   (rhoR nu_R dR v - rhoL nu_L dL v)/(2h) into m_f, the same
#                 with the face-mean m_f into e    hyperdifmomsource1.cu:17-533
#   m off-diag.:  d_f[rho (nu_R+nu_L)/4 d_d v_f] into m_d, with m_d for rho
#                 into e                           hyperdifmomsourcene1.cu:379-463
#   b diagonal:   (nu_R dR b_f - nu_L dL b_f)/h into b_f, face-mean b_f into e
#                                                  hyperdifbsource1.cu:21-440
#   b off-diag.:  -d_f[(nu_R+nu_L)/2 d_d b_f] into b_d, times b_d into e
#                                                  hyperdifbsourcene1.cu:18-820
# with the (ii, ii0) and (jj, ii0, mm, sb) patterns of iosmaug.cpp:1046-1144.
# Every term is compared cell by cell, so a flipped sign, a changed factor,
# a dropped heating term, a transposed derivative or a wrong cmax/dx/chyp
# index fails one of them.  dR, dL are one-sided differences /h and d the
# centred difference (grad13n) /2h; both are exact on the polynomials used.

A3 = 2.0 ** -10           # cubic amplitude: every value below is exact in FP64
BOX_N = (14, 13, 12)
BOX_H = (0.7, 1.3, 0.9)
BOX_CMAX = np.array([1.7, 2.3, 0.6])
CHI = 0.3
RHOB = 2.0


def box(n=BOX_N, h=BOX_H, bc=(4, 4, 4)):
    nc = int(np.prod(n))
    s = configs.Setup(name="box", ndim=3, n=tuple(n), bc=bc, gamma=5.0 / 3.0,
                      g=(0.0, 0.0, 0.0), chyp=np.zeros(13), hyperdif=1, driver=0,
                      drv=np.zeros(8), dx=[np.full(n[d], h[d]) for d in range(3)],
                      pos=[np.arange(n[d]) * h[d] for d in range(3)],
                      w=np.zeros((8, nc)), bg=np.zeros((5, nc)))
    s.bg[0] = 10.0
    s.bg[1] = RHOB
    return s


def hyper_part(s, w, lib_path=None):
    """L(w) with hyperdiffusion minus L(w) without it, at the fixed cmax"""
    o = Oracle(s, lib_path)
    o.g.hyperdif = 1
    on = o.rhs(w, cmax=BOX_CMAX)
    o.g.hyperdif = 0
    off = o.rhs(w, cmax=BOX_CMAX)
    return on - off


def _D1(f):
    return 3.0 * f * f + 3.0 * f + 1.0


def _cubic_terms(x, d, chi):
    """nu_R, nu_L and the one-sided differences dR, dL (per unit amplitude)
    of q = x^3 along d at index x"""
    h = BOX_H[d]
    K = h * BOX_CMAX[d] * chi
    nuR = K * 6.0 / _D1(x + 2)
    nuL = K * 6.0 / _D1(x + 1)
    dR = _D1(x) / h
    dL = _D1(x - 1) / h
    return nuR, nuL, dR, dL


def _core(s, dims):
    """cells whose index along each dim in ``dims`` lies in [3, n-4] (no pad
    and no edge rule reaches them) and that are interior along the others"""
    xi = cell_coords(s)
    m = interior_mask(s)
    for d in dims:
        m = m & (xi[d] >= 3) & (xi[d] <= s.n[d] - 4)
    return m


def _close(got, expect, m, what):
    scale = np.abs(expect[m]).max()
    assert scale > 0, what
    err = np.abs(got[m] - expect[m]).max()
    assert err <= 1e-11 * scale, f"{what}: max error {err:.3e} of {scale:.3e}"


def _zero(hp, fields, m, what):
    for f in fields:
        assert np.abs(hp[f][m]).max() <= 1e-14 * (1 + np.abs(hp).max()), f"{what}: field {f}"


def check_rho_and_energy_sources(lib_path=None):
    for fld in (0, 4):
        for d in range(3):
            s = box()
            s.chyp[fld] = CHI
            x = cell_coords(s)[d].astype(float)
            w = np.zeros((8, s.ncell))
            w[fld] = A3 * x ** 3 + (5.0 if fld == 4 else 0.0)
            hp = hyper_part(s, w, lib_path)
            nuR, nuL, dR, dL = _cubic_terms(x, d, CHI)
            expect = A3 * (nuR * dR - nuL * dL) / BOX_H[d]
            m = _core(s, [d])
            _close(hp[fld], expect, m, f"field {fld} along {d}")
            _zero(hp, [f for f in range(8) if f != fld], m, f"field {fld} along {d}")


def check_momentum_diagonal_sources(lib_path=None):
    # v_d = m_d/rho = a x_d^3 with rho = rho_b: the (v_d, d) pair is diagonal
    # in both calls of the (ii, ii0) loop, so m_d gets it twice
    for d in range(3):
        s = box()
        s.chyp[1 + d] = CHI
        x = cell_coords(s)[d].astype(float)
        w = np.zeros((8, s.ncell))
        w[1 + d] = RHOB * A3 * x ** 3
        hp = hyper_part(s, w, lib_path)
        nuR, nuL, dR, dL = _cubic_terms(x, d, CHI)
        h = BOX_H[d]
        m = _core(s, [d])
        em = 2 * 0.5 * A3 * (RHOB * nuR * dR - RHOB * nuL * dL) / h
        mR = RHOB * A3 * (x ** 3 + (x + 1) ** 3) / 2
        mL = RHOB * A3 * (x ** 3 + (x - 1) ** 3) / 2
        ee = 2 * 0.5 * A3 * (mR * nuR * dR - mL * nuL * dL) / h
        _close(hp[1 + d], em, m, f"m{d} diagonal")
        _close(hp[4], ee, m, f"m{d} diagonal heating")
        _zero(hp, [f for f in range(8) if f not in (1 + d, 4)], m, f"m{d} diagonal")


def check_momentum_offdiagonal_sources(lib_path=None):
    # v_f = a x_d^3 exactly (m_f = rho v_f, rho = rho_b + r x_f), m_d = s x_f + m0
    r, sl, m0 = 2.0 ** -6, 2.0 ** -5, 0.125
    for d in range(3):
        for f in range(3):
            if f == d:
                continue
            s = box()
            s.chyp[1 + f] = CHI
            xi = cell_coords(s)
            x, y = xi[d].astype(float), xi[f].astype(float)
            rho = RHOB + r * y
            w = np.zeros((8, s.ncell))
            w[0] = r * y
            w[1 + f] = rho * A3 * x ** 3
            w[1 + d] = sl * y + m0
            hp = hyper_part(s, w, lib_path)
            nuR, nuL, dR, dL = _cubic_terms(x, d, CHI)
            hd, hf = BOX_H[d], BOX_H[f]
            m = _core(s, [d, f])
            # diagonal (v_f, d) into m_f, heating with the face-mean m_f
            emf = 0.5 * A3 * rho * (nuR * dR - nuL * dL) / hd
            mR = rho * A3 * (x ** 3 + (x + 1) ** 3) / 2
            mL = rho * A3 * (x ** 3 + (x - 1) ** 3) / 2
            ediag = 0.5 * A3 * (mR * nuR * dR - mL * nuL * dL) / hd
            # off-diagonal: d_f[rho g6] into m_d, d_f[g6 m_d] into e, with
            # g6 = d_d v_f (nu_R + nu_L)/4 independent of x_f
            g6 = A3 * (6 * x * x + 2) / (2 * hd) * (nuR + nuL) / 4
            emd = g6 * r / hf
            eoff = g6 * sl / hf
            _close(hp[1 + f], emf, m, f"(v{f}, dim {d}) diagonal into m{f}")
            _close(hp[1 + d], emd, m, f"(v{f}, dim {d}) off-diagonal into m{d}")
            _close(hp[4], ediag + eoff, m, f"(v{f}, dim {d}) heating")
            t = 3 - d - f
            _zero(hp, [0, 1 + t, 5, 6, 7], m, f"(v{f}, dim {d})")


def check_magnetic_sources(lib_path=None):
    # b_f = a x_d^3 (1 + r x_f), b_d = beta: nu of (b_f, d) is the cubic's,
    # d_d b_f is linear in x_f
    r, beta = 2.0 ** -4, 0.75
    for d in range(3):
        for f in range(3):
            if f == d:
                continue
            s = box()
            s.chyp[5 + f] = CHI
            xi = cell_coords(s)
            x, y = xi[d].astype(float), xi[f].astype(float)
            fac = 1.0 + r * y
            w = np.zeros((8, s.ncell))
            w[5 + f] = A3 * x ** 3 * fac
            w[5 + d] = beta
            hp = hyper_part(s, w, lib_path)
            nuR, nuL, dR, dL = _cubic_terms(x, d, CHI)
            hd, hf = BOX_H[d], BOX_H[f]
            m = _core(s, [d, f])
            ebf = A3 * fac * (nuR * dR - nuL * dL) / hd
            bR = A3 * fac * (x ** 3 + (x + 1) ** 3) / 2
            bL = A3 * fac * (x ** 3 + (x - 1) ** 3) / 2
            ediag = A3 * fac * (bR * nuR * dR - bL * nuL * dL) / hd
            ebd = -(nuR + nuL) / 2 * A3 * (6 * x * x + 2) / (2 * hd) * r / hf
            _close(hp[5 + f], ebf, m, f"(b{f}, dim {d}) diagonal into b{f}")
            _close(hp[5 + d], ebd, m, f"(b{f}, dim {d}) off-diagonal into b{d}")
            _close(hp[4], ediag + beta * ebd, m, f"(b{f}, dim {d}) heating")
            t = 3 - d - f
            _zero(hp, [0, 1, 2, 3, 5 + t], m, f"(b{f}, dim {d})")


def check_hyperdiffusion_dissipates(lib_path=None):
    """on a rough periodic field each diagonal hyperdiffusion term alone
    lowers sum q^2 (sum of q dq/dt < 0) for q = rho', e', v_f and b_f"""
    rng = np.random.default_rng(17)
    for fld in (0, 4, 1, 2, 3, 5, 6, 7):
        s = box(n=(12, 12, 12), bc=(0, 0, 0))
        s.chyp[fld] = CHI
        w = np.zeros((8, s.ncell))
        w[fld] = 1e-2 * rng.standard_normal(s.ncell) + (5.0 if fld == 4 else 0.0)
        w = Oracle(s, lib_path).boundary(w)
        hp = hyper_part(s, w, lib_path)
        m = interior_mask(s)
        q = w[fld] / RHOB if 1 <= fld <= 3 else w[fld] - (5.0 if fld == 4 else 0.0)
        prod = (q * hp[fld])[m]
        assert prod.sum() < -0.5 * np.abs(prod).sum(), f"field {fld} is not dissipated"


def check_maxcourant_stretched(lib_path=None):
    """uniform state on a grid stretched along one dim: the Courant limit is
    cf_d / min(dx_d) along the dim that maximises it (each cell divides by
    its own dx; the smallest cell is not at the edge)"""
    gam = 5.0 / 3.0
    rho, p = 1.3, 0.7
    v = np.array([0.2, -0.4, 0.1])
    B = np.array([0.5, 0.3, -0.6])
    cs2 = gam * p / rho
    va2 = B @ B / rho
    cf = [math.sqrt(0.5 * (cs2 + va2 + math.sqrt((cs2 + va2) ** 2 - 4 * cs2 * B[d] ** 2 / rho)))
          + abs(v[d]) for d in range(3)]
    for sd in range(3):
        s = box(n=(12, 11, 10))
        dxs = []
        for d in range(3):
            n = s.n[d]
            if d == sd:
                t = np.linspace(-1.0, 1.0, n)
                dxs.append(0.05 * (1.0 + 3.0 * t * t + 0.5 * t))  # minimum inside
            else:
                dxs.append(np.full(n, 1.0))
        s.dx = dxs
        s.bg[:] = 0.0
        w = np.zeros((8, s.ncell))
        w[0] = rho
        w[1:4] = (rho * v)[:, None]
        w[4] = p / (gam - 1) + 0.5 * rho * v @ v + 0.5 * B @ B
        w[5:8] = B[:, None]
        got = Oracle(s, lib_path).maxcourant(w)
        expect = cf[sd] / dxs[sd].min()
        assert abs(got - expect) <= 1e-12 * expect, (sd, got, expect)


def check_numax_and_dtvisc(lib_path=None):
    """numax: a field linear along d has D3 = 0 except where a window reaches
    a pad, where nu = h_d cmax_d chyp 6/5 (q[-1] = q[5] gives D3 = 6,
    max D1 = 5); no other pair of the state has a nonzero coefficient, and a
    b_f that varies only along f has none at all.  dtvisc: 0.25 min dx^2 /
    numax over the dims and the cells' own dx."""
    chyp = np.zeros(13)
    chyp[[0, 1, 2, 3, 4, 5, 6, 7]] = [0.11, 0.23, 0.29, 0.31, 0.17, 0.13, 0.19, 0.37]
    for fld in range(8):
        for d in range(3):
            s = box()
            s.chyp = chyp.copy()
            x = cell_coords(s)[d].astype(float)
            w = np.zeros((8, s.ncell))
            w[fld] = (RHOB * x) if 1 <= fld <= 3 else x + (5.0 if fld == 4 else 0.0)
            # e' cancels the kinetic or magnetic part of the internal energy
            # e' - (b'^2 + m^2/rho)/2 (exact on these integers), so only
            # field fld carries a coefficient
            if 1 <= fld <= 3:
                w[4] = w[fld] ** 2 / (2 * RHOB)
            elif fld >= 5:
                w[4] = w[fld] ** 2 / 2
            got = Oracle(s, lib_path).numax(w, cmax=BOX_CMAX)
            if fld >= 5 and fld - 5 == d:
                expect = 0.0  # (b_d, d) is not a hyperdiffusion pair
            else:
                expect = BOX_H[d] * BOX_CMAX[d] * chyp[fld] * 6.0 / 5.0
            assert abs(got - expect) <= 1e-14 * max(expect, 1e-300), (fld, d, got, expect)
    # one-sided profiles: linear at one end and flat at the other, so the
    # pad value 6/5 comes from nu_L at x = 2 alone (q = min(x, 5)) or from
    # nu_R at x = n-3 alone (q = max(x, n-6)); the kink gives 1
    for d in range(3):
        s = box()
        s.chyp = chyp.copy()
        x = cell_coords(s)[d].astype(float)
        n = s.n[d]
        for q in (np.minimum(x, 5.0), np.maximum(x, n - 6.0)):
            w = np.zeros((8, s.ncell))
            w[0] = q
            got = Oracle(s, lib_path).numax(w, cmax=BOX_CMAX)
            expect = BOX_H[d] * BOX_CMAX[d] * chyp[0] * 6.0 / 5.0
            assert abs(got - expect) <= 1e-14 * expect, (d, got, expect)
    s = box(n=(12, 11, 10))
    for d in range(3):
        t = np.linspace(-1.0, 1.0, s.n[d])
        s.dx[d] = (0.3 + d * 0.1) * (1.0 + 2.0 * t * t + 0.3 * t)
    o = Oracle(s, lib_path)
    numax = 0.37
    expect = 0.25 * min(float((dx ** 2).min()) for dx in s.dx) / numax
    got = o.dtvisc(numax)
    assert abs(got - expect) <= 1e-15 * expect, (got, expect)
    assert o.dtvisc(0.0) == 0.0


def check_cont_and_fixed_ghosts(lib_path=None):
    """cont (3): both ghost layers of a face copy its edge cell (2, n-3);
    fixed (5): the ghosts take the captured state's values.  Dim 0 is filled
    first over the whole transverse range, then dim 1, so the corners follow
    dim 1's rule."""
    s = grid2d(nx=10, ny=9, bc=(3, 5))
    i, j, _ = cell_coords(s)
    n0, n1 = s.n[0], s.n[1]
    f = np.sin(0.7 * i) + np.cos(0.3 * j) + 2.0
    w = np.zeros((s.nve, s.ncell))
    w[0] = np.where(interior_mask(s), f, 1e9)
    fix = np.zeros_like(w)
    fix[0] = 1000.0 + i + 100.0 * j
    o = Oracle(s, lib_path)
    o.set_fixed(fix)
    got = o.boundary(w)[0].reshape(n1, n0)
    F = f.reshape(n1, n0)
    rows = slice(2, n1 - 2)
    assert np.array_equal(got[rows, 0], F[rows, 2]) and np.array_equal(got[rows, 1], F[rows, 2])
    assert np.array_equal(got[rows, n0 - 1], F[rows, n0 - 3])
    assert np.array_equal(got[rows, n0 - 2], F[rows, n0 - 3])
    fx = fix[0].reshape(n1, n0)
    for jj in (0, 1, n1 - 2, n1 - 1):
        assert np.array_equal(got[jj], fx[jj])
    assert np.array_equal(got[2:n1 - 2, 2:n0 - 2], F[2:n1 - 2, 2:n0 - 2])


SOURCE_CHECKS = [check_rho_and_energy_sources, check_momentum_diagonal_sources,
                 check_momentum_offdiagonal_sources, check_magnetic_sources,
                 check_hyperdiffusion_dissipates, check_maxcourant_stretched,
                 check_numax_and_dtvisc, check_cont_and_fixed_ghosts]


@pytest.mark.parametrize("check", SOURCE_CHECKS, ids=lambda c: c.__name__[6:])
def test_hyperdiffusive_sources_and_courant(check):
    check()
