"""Launch-shape tuning entry points: the autotuner and the timing calls leave
the state alone, every shape they can pick gives the same L(w) as the
default one, and the reported pass times are sane."""
import numpy as np
import pytest

from helpers import interior_mask, perturbed, rel_err

pytestmark = pytest.mark.gpu

from smaugamd import configs  # noqa: E402
from smaugamd.engine import FP64, MIXED, Engine  # noqa: E402
from oracle.oracle import Oracle  # noqa: E402


def strat():
    return configs.stratified(nz=36, nx=20, ny=12, lz=6e6 * 36 / 256 * 4,
                              lx=4e6 * 20 / 256 * 4, ly=4e6 * 12 / 256 * 4)


@pytest.mark.parametrize("prec", [MIXED, FP64], ids=["mixed", "fp64"])
def test_autotune_keeps_state_and_operator(prec):
    s = strat()
    print(f"[case] autotune prec={prec} n={s.n}", flush=True)
    o = Oracle(s)
    w = o.boundary(perturbed(s, amp=1e-2, seed=9))
    e = Engine(s, prec)
    e.set_state(w)
    ref = e.rhs(t=0.5)
    before = e.get_state()
    best, ms = e.autotune(1)
    print("autotune", best, ms, e.get_zstream())
    assert np.array_equal(e.get_state(), before)
    assert ms[1] > 0
    got = e.rhs(t=0.5)
    tol = 1e-12 if prec == FP64 else 2e-5
    assert np.all(rel_err(got, ref, interior_mask(s)) < tol)
    # pass times: five passes, non-negative, the stage update the largest
    # share of a step together with the sources
    p = e.time_passes(0.0, 0.1 * configs.hyperviscous_dt(s))
    print("passes", p)
    assert len(p) == 5 and all(v >= 0 for v in p.values())
    e.close()


def test_panel_and_forms_roundtrip():
    s = strat()
    print("[case] panel/form setters", flush=True)
    e = Engine(s, MIXED)
    e.set_state()
    ntx, nty = (s.n[0] - 4 + 63) // 64, (s.n[1] - 4 + 3) // 4
    e.set_panel(1, 2, 2)  # panels are clamped to the tile counts
    assert e.get_panel(1) == (min(2, ntx), min(2, nty))
    e.set_panel(1, 64, 64)
    assert e.get_panel(1) == (ntx, nty)
    for form in (0, 1, 2, 3):
        e.set_zstream(form, 5)
        assert e.get_zstream()[0] == form
    e.set_ratio_form(False)
    e.set_ratio_form(True, 7)
    assert e.time_pass(0, 1) > 0 and e.time_pass(1, 1) > 0
    e.close()
