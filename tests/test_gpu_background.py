"""Background layouts of the per-cell kernels: a background that is the same
in every cell is read as its few values, any other in full.  Both against
the oracle on split states (part of the state moved into the background),
and the ensemble, whose runs choose their layout one by one, against the
single-grid engine bit for bit."""
import dataclasses

import numpy as np
import pytest

from helpers import interior_mask, perturbed, rel_err

pytestmark = pytest.mark.gpu

from smaugamd import configs  # noqa: E402
from smaugamd.engine import FP64, MIXED, Engine, Ensemble  # noqa: E402
from oracle.oracle import Oracle  # noqa: E402

TOL_FP64 = 1e-10


def split(setup, w, uniform):
    """the same total state with part of rho, e and b moved into the
    background: constants (uniform) or half of the local values"""
    D = setup.ndim
    bg = np.array(setup.bg, copy=True)
    w = np.array(w, copy=True)
    parts = [(1, 0), (0, 1 + D)] + [(2 + d, 2 + D + d) for d in range(D)]  # (bg field, w field)
    for fb, fw in parts:
        share = (0.3 * np.mean(w[fw]) + 0.05) * np.ones_like(w[fw]) if uniform else 0.5 * w[fw]
        bg[fb] += share
        w[fw] -= share
    return dataclasses.replace(setup, bg=bg), w


@pytest.mark.parametrize("uniform", [True, False], ids=["uniform", "varying"])
def test_split_background_vs_oracle(uniform):
    base = configs.orszag_tang(nx=48)
    o0 = Oracle(base)
    w0 = o0.boundary(perturbed(base, amp=1e-2, seed=41))
    s, w = split(base, w0, uniform)
    print(f"[case] split background uniform={uniform} n={s.n}", flush=True)
    o = Oracle(s)
    w = o.boundary(w)
    ref_r = o.rhs(w, t=0.5)
    ref, _ = o.run(w, 0.5, s.dt, 3, rk4=True)
    e = Engine(s, FP64)
    e.set_state(w)
    got_r = e.rhs(t=0.5)
    e.rk4(0.5, s.dt, 3)
    got = e.get_state()
    e.close()
    m = interior_mask(s)
    err = rel_err(got_r, ref_r, m)
    assert np.all(err < TOL_FP64), err
    err = rel_err(got, ref, m)
    assert np.all(err < TOL_FP64), err


@pytest.mark.parametrize("prec", [FP64, MIXED], ids=["fp64", "mixed"])
def test_ensemble_mixed_layouts_match_single_grid(prec):
    a = configs.alfven_member(40, 1.1)
    b = configs.orszag_tang(nx=56, ny=36)
    b.dt = 0.3 / configs.initial_max_courant(b)
    sa, wa = split(a, Oracle(a).boundary(perturbed(a, amp=1e-2, seed=42)), True)
    sb, wb = split(b, Oracle(b).boundary(perturbed(b, amp=1e-2, seed=43)), False)
    setups, ws = [sa, sb], [Oracle(sa).boundary(wa), Oracle(sb).boundary(wb)]
    print(f"[case] ensemble uniform + varying backgrounds prec={prec}", flush=True)
    dts = [s.dt for s in setups]
    ens = Ensemble(setups, prec)
    for r, w in enumerate(ws):
        ens.set_state(r, w)
    ens.rk4([0.0, 0.0], dts, 4)
    for r, s in enumerate(setups):
        e = Engine(s, prec)
        e.set_state(ws[r])
        e.rk4(0.0, dts[r], 4)
        ref = e.get_state()
        e.close()
        got = ens.get_state(r)
        assert np.all(np.isfinite(got))
        assert np.array_equal(got, ref), rel_err(got, ref)
    ens.close()
