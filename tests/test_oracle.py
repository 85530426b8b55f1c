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
