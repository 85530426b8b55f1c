"""2-D ratio form 2: the stage update forms the hyperviscosity ratios of its
tile and a 1-cell halo in LDS (k_update2r) instead of reading them from R,
which the ratio pass (k_ratio) would fill.  Every ratio is the same ratio6 of
the same six values and the update arithmetic is update_cell's, so L(w), the
RK4 steps (the folded aux pass of stages 0-2 and stage 3) and the Euler step
agree with form 0 bit for bit -- on ragged tiles, the smallest grid, a
uniform and a non-uniform background -- and with the oracle in FP64."""
import dataclasses

import numpy as np
import pytest

from helpers import perturbed, rel_err

pytestmark = pytest.mark.gpu

from smaugamd import configs  # noqa: E402
from smaugamd.engine import FP64, MIXED, Engine  # noqa: E402
from oracle.oracle import Oracle  # noqa: E402


def bw_bg(ncells):
    """Brio-Wu with half of the density as a background that varies along x"""
    s = configs.brio_wu(ncells=ncells)
    w, bg = s.w.copy(), s.bg.copy()
    bg[1] = 0.5 * w[0]
    w[0] = w[0] - bg[1]
    return dataclasses.replace(s, w=w, bg=bg)


CASES = [
    ("ot32", configs.orszag_tang(nx=32)),
    ("ot70x9", configs.orszag_tang(nx=70, ny=9)),     # two i-tiles, ragged j-tile
    ("ot130x7", configs.orszag_tang(nx=130, ny=7)),   # three i-tiles
    ("ot4", configs.orszag_tang(nx=4)),               # the smallest grid
    ("bw64", configs.brio_wu(ncells=64)),             # contcd4 faces, 4 rows
    ("bw64_bg", bw_bg(64)),                           # background read per cell
    ("alfven96", configs.alfven_member(96, 1.1)),
]


def _run(s, prec, form, w0, dt):
    e = Engine(s, prec)
    e.set_ratio_form(form)
    assert e.get_ratio_form()[0] == form
    e.set_state(w0, s.bg)
    rhs = e.rhs(t=0.5)
    e.rk4(0.25, dt, 3)
    out = e.get_state()
    e.set_state(w0, s.bg)
    e.euler(0.25, dt, 2)
    eul = e.get_state()
    e.close()
    return rhs, out, eul


@pytest.mark.parametrize("prec", [FP64, MIXED], ids=["fp64", "mixed"])
@pytest.mark.parametrize("name,setup", CASES, ids=[c[0] for c in CASES])
def test_ratio_form2_2d_matches_ratio_pass(name, setup, prec):
    print(f"[case] 2-D ratio form 2 vs 0 {name} n={setup.n} prec={prec}", flush=True)
    o = Oracle(setup)
    w0 = o.boundary(perturbed(setup, amp=1e-2, seed=31))
    if prec == MIXED:
        w0 = o.boundary(w0.astype(np.float32).astype(np.float64))
    dt = 0.5 * configs.hyperviscous_dt(setup)
    r0, s0, e0 = _run(setup, prec, 0, w0, dt)
    r2, s2, e2 = _run(setup, prec, 2, w0, dt)
    assert np.all(np.isfinite(r2)) and np.all(np.isfinite(s2))
    assert np.array_equal(r2, r0), rel_err(r2, r0)
    assert np.array_equal(s2, s0), rel_err(s2, s0)
    assert np.array_equal(e2, e0), rel_err(e2, e0)


@pytest.mark.parametrize("name,setup", [CASES[1], CASES[5]], ids=[CASES[1][0], CASES[5][0]])
def test_ratio_form2_2d_rk4_fp64_vs_oracle(name, setup):
    print(f"[case] 2-D ratio form 2 rk4 vs oracle {name} n={setup.n}", flush=True)
    o = Oracle(setup)
    w0 = o.boundary(perturbed(setup, amp=1e-3, seed=32))
    dt = 0.5 * configs.hyperviscous_dt(setup)
    ref, _ = o.run(w0, 0.25, dt, 3, rk4=True)
    _, got, _ = _run(setup, FP64, 2, w0, dt)
    err = rel_err(got, ref)
    print(name, "2-D form 2 rk4 rel err vs oracle", err)
    assert np.all(err < 1e-10), err


def test_ratio_form1_rejected_in_2d():
    s = CASES[0][1]
    e = Engine(s, MIXED)
    from smaugamd.engine import SmaugError
    with pytest.raises(SmaugError, match="z-streamed ratios are for 3-D grids"):
        e.set_ratio_form(1)
    e.close()
