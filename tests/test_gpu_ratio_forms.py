"""Ratio form 2: the ratio pass forms the seven hyperviscosity-ratio pairs
along k alone (k_ratio_kz) and the z-streamed source kernel forms the
fourteen in-plane pairs in LDS from the state around its tile
(k_hyper_zsr).  Every ratio is the same ratio6 of the same six values as in
form 1 (k_ratio_zs), and the source kernel sums the terms in the same order,
so L(w) and the RK4 steps agree with form 1 bit for bit -- on ragged tiles,
short z-chunks, periodic faces, a k-varying background, and the launch
shapes of the timed runs -- and with the oracle in FP64."""
import dataclasses

import numpy as np
import pytest

from helpers import perturbed, rel_err

pytestmark = pytest.mark.gpu

from smaugamd import configs  # noqa: E402
from smaugamd.engine import FP64, MIXED, Engine  # noqa: E402
from oracle.oracle import Oracle  # noqa: E402


def strat(nz, nx, ny):
    return configs.stratified(nz=nz, nx=nx, ny=ny, lz=6e6 * nz / 256 * 4,
                              lx=4e6 * nx / 256 * 4, ly=4e6 * ny / 256 * 4)


def uniform_bg(s):
    """background of one interior cell everywhere, no gravity: periodic
    along every dim"""
    bg = np.repeat(s.bg[:, [s.n[0] * 3 + 2]], s.ncell, axis=1)
    return dataclasses.replace(s, bg=bg, g=(0.0, 0.0, 0.0))


def k_varying(s):
    bg = s.bg.copy()
    n0, n1, n2 = s.n
    kk = np.arange(n2).reshape(n2, 1, 1)
    bg[1] *= (1.0 + 1e-3 * np.sin(kk)).repeat(n1, 1).repeat(n0, 2).reshape(-1)
    return dataclasses.replace(s, bg=bg)


# (name, set-up, update/source depth, ratio depth)
CASES = [
    ("70x9x13", strat(70, 9, 13), 0, 0),            # two i-tiles, ragged j-tile
    ("20x12x16_kz5_r3", strat(20, 12, 16), 5, 3),   # short chunks along k
    ("24x10x12_kz1_r1", strat(24, 10, 12), 1, 1),   # one plane per block
    ("130x7x9", strat(130, 7, 9), 4, 2),            # three i-tiles, ragged j-tile
    ("ppp_uniform", dataclasses.replace(uniform_bg(strat(20, 12, 16)), bc=(0, 0, 0)), 4, 4),
    ("pcc_uniform", dataclasses.replace(uniform_bg(strat(20, 12, 16)), bc=(0, 4, 4)), 6, 5),
    ("cpp", dataclasses.replace(strat(20, 12, 16), bc=(4, 0, 0)), 8, 8),
    ("k_varying_bg", k_varying(strat(20, 12, 16)), 5, 4),
    ("timed_kz128_r8", strat(12, 12, 260), 128, 8),
]


def _run(s, prec, form, kz, kzr, w0, dt):
    e = Engine(s, prec)
    e.set_zstream(True, kz)
    e.set_ratio_form(form, kzr)
    assert e.get_ratio_form()[0] == form
    e.set_state(w0, s.bg)
    rhs = e.rhs(t=0.5)
    e.rk4(0.25, dt, 3)
    out = e.get_state()
    e.close()
    return rhs, out


@pytest.mark.parametrize("prec", [FP64, MIXED], ids=["fp64", "mixed"])
@pytest.mark.parametrize("name,setup,kz,kzr", CASES, ids=[c[0] for c in CASES])
def test_ratio_form2_matches_form1(name, setup, kz, kzr, prec):
    print(f"[case] ratio form 2 vs 1 {name} n={setup.n} kz={kz} kzr={kzr} prec={prec}",
          flush=True)
    o = Oracle(setup)
    w0 = o.boundary(perturbed(setup, amp=1e-2, seed=21))
    if prec == MIXED:
        w0 = o.boundary(w0.astype(np.float32).astype(np.float64))
    dt = 0.5 * configs.hyperviscous_dt(setup)
    r1, s1 = _run(setup, prec, 1, kz, kzr, w0, dt)
    r2, s2 = _run(setup, prec, 2, kz, kzr, w0, dt)
    assert np.all(np.isfinite(r2)) and np.all(np.isfinite(s2))
    assert np.array_equal(r2, r1), rel_err(r2, r1)
    assert np.array_equal(s2, s1), rel_err(s2, s1)


@pytest.mark.parametrize("name,setup,kz,kzr", [CASES[0], CASES[4], CASES[8]],
                         ids=[CASES[0][0], CASES[4][0], CASES[8][0]])
def test_ratio_form2_rk4_fp64_vs_oracle(name, setup, kz, kzr):
    print(f"[case] ratio form 2 rk4 vs oracle {name} n={setup.n}", flush=True)
    o = Oracle(setup)
    w0 = o.boundary(perturbed(setup, amp=1e-3, seed=22))
    dt = 0.5 * configs.hyperviscous_dt(setup)
    ref, _ = o.run(w0, 0.25, dt, 3, rk4=True)
    _, got = _run(setup, FP64, 2, kz, kzr, w0, dt)
    err = rel_err(got, ref)
    print(name, "form 2 rk4 rel err vs oracle", err)
    assert np.all(err < 1e-10), err


def test_ratio_form2_needs_zstreamed_sources():
    # with the per-cell source kernel the ratio pass forms every pair again
    name, s, kz, kzr = CASES[1]
    print(f"[case] ratio form 2 with per-cell sources {name}", flush=True)
    o = Oracle(s)
    w0 = o.boundary(perturbed(s, amp=1e-3, seed=23))
    e = Engine(s, FP64)
    e.set_state(w0, s.bg)
    e.set_zstream(2, kz)  # z-streamed update, per-cell source kernel
    e.set_ratio_form(2, kzr)
    got = e.rhs(t=0.5)
    e.close()
    ref = o.rhs(w0, 0.5)
    from helpers import interior_mask
    err = rel_err(got, ref, interior_mask(s))
    assert np.all(err < 1e-10), err
