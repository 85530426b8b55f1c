"""The ensemble (smaug_ens_* in smaug_engine.h): uneven grids advanced by the
persistent queue-driven kernels evolve like the same grids stepped alone,
and in FP64 like the oracle."""
import numpy as np
import pytest

from helpers import interior_mask, perturbed, rel_err

from smaugamd import configs
from smaugamd.engine import FP64, MIXED, Engine, Ensemble, SmaugError
from oracle.oracle import Oracle

# an ensemble member and the single-grid engine run the same per-cell device
# code on the same inputs, so their states agree bit for bit
TOL_ORACLE_FP64 = 1e-11


def members_2d():
    # uneven sides, one non-square grid, per-run dt at 0.3 CFL
    out = [configs.alfven_member(24, 0.7), configs.alfven_member(40, 1.3),
           configs.orszag_tang(nx=72, ny=20)]
    out[2].dt = 0.3 / configs.initial_max_courant(out[2])
    return out


def members_3d():
    def strat(nz, nx, ny):
        return configs.stratified(nz=nz, nx=nx, ny=ny, lz=6e6 * nz / 256 * 4,
                                  lx=4e6 * 16 / 256, ly=4e6 * 16 / 256)
    return [strat(20, 12, 16), strat(28, 16, 10)]


def states(setups, amp, seed):
    out = []
    for k, s in enumerate(setups):
        o = Oracle(s)
        out.append(o.boundary(perturbed(s, amp=amp, seed=seed + k)))
    return out


def test_mismatched_physics_rejected():
    # rejected before any device work, so this runs without a GPU
    a = configs.alfven_member(24, 1.0)
    b = configs.alfven_member(32, 1.0)
    b.gamma = 1.4
    with pytest.raises(SmaugError, match="must be the same"):
        Ensemble([a, b], MIXED)


@pytest.mark.gpu
@pytest.mark.parametrize("prec", [MIXED, FP64], ids=["mixed", "fp64"])
@pytest.mark.parametrize("dim", [2, 3])
def test_member_matches_single_grid(dim, prec):
    setups = members_2d() if dim == 2 else members_3d()
    print(f"[case] ensemble vs single grid dim={dim} prec={prec} "
          f"sizes={[s.n for s in setups]}", flush=True)
    amp = 1e-2 if dim == 2 else 1e-3
    w0 = states(setups, amp, seed=11)
    dts = [s.dt if s.dt > 0 else 0.3 / configs.initial_max_courant(s) for s in setups]
    if dim == 3:
        dts = [0.5 * configs.hyperviscous_dt(s) for s in setups]
    t0 = [0.0, 1.5, 3.0][:len(setups)]
    ens = Ensemble(setups, prec)
    for r, w in enumerate(w0):
        ens.set_state(r, w)
    nsteps = 6
    ens.rk4(t0, dts, nsteps)
    for r, s in enumerate(setups):
        e = Engine(s, prec)
        if dim == 3:
            e.set_zstream(False)  # the ensemble runs the per-cell stage update
        e.set_state(w0[r])
        e.rk4(t0[r], dts[r], nsteps)
        ref = e.get_state()
        got = ens.get_state(r)
        e.close()
        same = np.array_equal(got, ref)
        print(f"run {r} n={s.n} bitwise={same} rel err {rel_err(got, ref, interior_mask(s))}")
        assert np.all(np.isfinite(got))
        assert same
    ens.close()


def periodic_bg_member(side):
    """an Orszag-Tang-type member with part of its density as a background
    that varies along x and is periodic (ghost cells copied from their
    images), so that the update reads it per cell"""
    s = configs.alfven_member(side, 0.9)
    n0, n1 = s.n[0], s.n[1]
    w, bg = s.w.copy(), s.bg.copy()
    i = np.arange(n0)
    f = 1.0 + 0.2 * np.sin(2 * np.pi * (i - 2) / (n0 - 4))
    rb = (0.5 * w[0].reshape(n1, n0)) * f[None, :]
    rb[:, 0:2] = rb[:, n0 - 4:n0 - 2]
    rb[:, n0 - 2:n0] = rb[:, 2:4]
    rb[0:2, :] = rb[n1 - 4:n1 - 2, :]
    rb[n1 - 2:n1, :] = rb[2:4, :]
    bg[1] = rb.reshape(-1)
    w[0] = w[0] - bg[1]
    s.w, s.bg = w, bg
    return s


@pytest.mark.gpu
@pytest.mark.parametrize("prec", [MIXED, FP64], ids=["mixed", "fp64"])
def test_ratio_form2_2d_ensemble_matches_ratio_pass(prec):
    # the update pulls 64 x 4 tiles and forms their ratios in LDS (form 2)
    # or reads those of the ratio pass (form 0): the same states bit for bit
    setups = members_2d() + [periodic_bg_member(70)]
    print(f"[case] ensemble ratio form 2 vs 0 prec={prec} sizes={[s.n for s in setups]}",
          flush=True)
    w0 = states(setups, 1e-2, seed=13)
    dts = [0.3 / configs.initial_max_courant(s) for s in setups]
    t0 = [0.0, 1.5, 3.0, 0.5]
    out = {}
    for form in (0, 2):
        ens = Ensemble(setups, prec)
        ens.set_ratio_form(form)
        assert ens.get_ratio_form() == form
        for r, w in enumerate(w0):
            ens.set_state(r, w, setups[r].bg)
        ens.rk4(t0, dts, 6)
        out[form] = [ens.get_state(r) for r in range(len(setups))]
        ens.close()
    for r, s in enumerate(setups):
        assert np.all(np.isfinite(out[2][r]))
        assert np.array_equal(out[2][r], out[0][r]), (r, rel_err(out[2][r], out[0][r]))


@pytest.mark.gpu
def test_fp64_ensemble_vs_oracle():
    setups = members_2d()[:2]
    print(f"[case] fp64 ensemble vs oracle sizes={[s.n for s in setups]}", flush=True)
    w0 = states(setups, 1e-2, seed=21)
    dts = [s.dt for s in setups]
    ens = Ensemble(setups, FP64)
    for r, w in enumerate(w0):
        ens.set_state(r, w)
    ens.rk4(0.0, dts, 4)
    for r, s in enumerate(setups):
        ref, _ = Oracle(s).run(w0[r], 0.0, dts[r], 4, rk4=True)
        err = rel_err(ens.get_state(r), ref, interior_mask(s))
        print(f"run {r} rel err vs oracle {err}")
        assert np.all(err < TOL_ORACLE_FP64), err
    ens.close()


@pytest.mark.gpu
def test_repeated_calls_continue_time():
    # two calls of 3 steps equal one call of 6 (the queues and cmax are
    # reset inside every replayed step)
    setups = members_2d()[:2]
    print("[case] ensemble 3+3 steps vs 6", flush=True)
    w0 = states(setups, 1e-2, seed=31)
    dts = [s.dt for s in setups]
    a = Ensemble(setups, MIXED)
    b = Ensemble(setups, MIXED)
    for r, w in enumerate(w0):
        a.set_state(r, w)
        b.set_state(r, w)
    a.rk4(0.0, dts, 6)
    b.rk4(0.0, dts, 3)
    b.rk4([3 * d for d in dts], dts, 3)
    for r in range(len(setups)):
        np.testing.assert_array_equal(a.get_state(r), b.get_state(r))
    a.close()
    b.close()


@pytest.mark.gpu
@pytest.mark.parametrize("dim", [2, 3])
def test_member_matches_single_grid_without_hyperdiffusion(dim):
    """without hyperdiffusion the ensemble holds no ratio buffer and runs no
    ratio pass; its members still evolve bit for bit like the single grid"""
    setups = members_2d() if dim == 2 else members_3d()
    for s in setups:
        s.hyperdif = 0
    print(f"[case] ensemble without hyperdiffusion dim={dim} sizes={[s.n for s in setups]}",
          flush=True)
    w0 = states(setups, 1e-2 if dim == 2 else 1e-3, seed=23)
    dts = [0.3 / configs.initial_max_courant(s) for s in setups]
    if dim == 3:
        dts = [0.5 * configs.hyperviscous_dt(s) for s in setups]
    t0 = [0.0, 1.5, 3.0][:len(setups)]
    ens = Ensemble(setups, MIXED)
    for r, w in enumerate(w0):
        ens.set_state(r, w)
    ens.rk4(t0, dts, 4)
    for r, s in enumerate(setups):
        e = Engine(s, MIXED)
        if dim == 3:
            e.set_zstream(False)
        e.set_state(w0[r])
        e.rk4(t0[r], dts[r], 4)
        ref = e.get_state()
        got = ens.get_state(r)
        e.close()
        print(f"run {r} n={s.n} bitwise={np.array_equal(got, ref)}")
        assert np.all(np.isfinite(got))
        assert np.array_equal(got, ref)
    ens.close()
