"""The z-streamed 3-D stage update (k_zstream) against the per-cell stage
update and against the FP64 oracle, on tiles and z-chunks that do not divide
the grid."""
import dataclasses

import numpy as np
import pytest

from helpers import interior_mask, perturbed, rel_err

pytestmark = pytest.mark.gpu

from smaugamd import configs  # noqa: E402
from smaugamd.engine import FP64, MIXED, Engine, SmaugError  # noqa: E402
from oracle.oracle import Oracle  # noqa: E402

# the two update kernels differ only in the summation order of the twelve
# neighbour contributions
TOL_RHS = {FP64: 1e-12, MIXED: 2e-5}
TOL_ORACLE_FP64 = 1e-10
# planar against full background after two RK4 steps (last-place roundings)
TOL_PLANE = {FP64: 1e-14, MIXED: 1e-6}


def strat(nz, nx, ny):
    return configs.stratified(nz=nz, nx=nx, ny=ny, lz=6e6 * nz / 256 * 4,
                              lx=4e6 * nx / 256 * 4, ly=4e6 * ny / 256 * 4)


CASES = [
    ("70x9x13", strat(70, 9, 13), 0),      # two i-tiles, ragged j-tile
    ("20x12x16_kz5", strat(20, 12, 16), 5),  # z-chunks of 5 over 16 planes
    ("64x8x8_kz3", strat(64, 8, 8), 3),
    # one plane per block: the most blocks and cfast partial maxima the
    # folded aux pass of the update can leave
    ("24x10x12_kz1", strat(24, 10, 12), 1),
]


@pytest.mark.parametrize("prec", [FP64, MIXED], ids=["fp64", "mixed"])
@pytest.mark.parametrize("name,setup,kz", CASES, ids=[c[0] for c in CASES])
def test_rhs_matches_per_cell_kernel(name, setup, kz, prec):
    print(f"[case] zstream rhs {name} kz={kz} prec={prec} n={setup.n}", flush=True)
    o = Oracle(setup)
    w = o.boundary(perturbed(setup, amp=1e-2, seed=5))
    if prec == MIXED:
        w = o.boundary(w.astype(np.float32).astype(np.float64))
    e = Engine(setup, prec)
    e.set_state(w)
    e.set_zstream(False)
    ref = e.rhs(t=2.0)
    # hyperdiffusive sources 1: z-streamed, 2: per cell, 3: plane staged in LDS
    for mode in (1, 2, 3):
        e.set_zstream(mode, kz)
        got = e.rhs(t=2.0)
        err = rel_err(got, ref, interior_mask(setup))
        print(name, f"zstream mode {mode} vs per-cell rel err", err)
        assert np.all(np.isfinite(got))
        assert np.all(err < TOL_RHS[prec]), err
        # ghosts stay untouched in the raw-L(w) mode
        assert np.all(got[:, ~interior_mask(setup)] == 0)
    e.close()


@pytest.mark.parametrize("name,setup,kz", CASES, ids=[c[0] for c in CASES])
def test_rk4_fp64_vs_oracle(name, setup, kz):
    print(f"[case] zstream rk4 fp64 vs oracle {name} kz={kz}", flush=True)
    o = Oracle(setup)
    w0 = o.boundary(perturbed(setup, amp=1e-3, seed=6))
    dt = 0.5 * configs.hyperviscous_dt(setup)
    ref, _ = o.run(w0, 0.25, dt, 4, rk4=True)
    e = Engine(setup, FP64)
    e.set_zstream(True, kz)
    e.set_state(w0)
    e.rk4(0.25, dt, 4)
    got = e.get_state()
    e.close()
    err = rel_err(got, ref, interior_mask(setup))
    print(name, "rk4 rel err vs oracle", err)
    assert np.all(err < TOL_ORACLE_FP64), err


@pytest.mark.parametrize("prec", [FP64, MIXED], ids=["fp64", "mixed"])
@pytest.mark.parametrize("name,setup,kz", CASES, ids=[c[0] for c in CASES])
def test_zstreamed_ratios_match_per_cell(name, setup, kz, prec):
    # the z-streamed ratio kernel evaluates the same windows with the same
    # arithmetic as the per-cell one, so L(w) is identical bit for bit
    print(f"[case] zstream ratios {name} kz={kz} prec={prec}", flush=True)
    o = Oracle(setup)
    w = o.boundary(perturbed(setup, amp=1e-2, seed=8))
    e = Engine(setup, prec)
    e.set_state(w)
    e.set_ratio_form(False)
    ref = e.rhs(t=1.0)
    e.set_ratio_form(True, kz)
    got = e.rhs(t=1.0)
    e.close()
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("prec", [FP64, MIXED], ids=["fp64", "mixed"])
@pytest.mark.parametrize("name,setup,kz", CASES, ids=[c[0] for c in CASES])
def test_bg_plane_matches_full_background(name, setup, kz, prec):
    # a stratified background is the same on every k-plane: the kernels then
    # read its one-plane copy, with the same arithmetic.  L(w) agrees bit for
    # bit with the full-background path; after two RK4 steps a few cells may
    # differ by a last-place rounding (the compiler may contract a
    # multiply-add differently in the two kernel instances)
    print(f"[case] bg plane {name} kz={kz} prec={prec}", flush=True)
    o = Oracle(setup)
    w0 = o.boundary(perturbed(setup, amp=1e-3, seed=9))
    dt = 0.5 * configs.hyperviscous_dt(setup)
    out = []
    for plane in (True, False):
        e = Engine(setup, prec)
        e.set_zstream(True, kz)
        e.set_bg_plane(plane)
        e.set_state(w0)
        assert e.get_bg_plane() == plane
        r = e.rhs(t=0.5)
        e.rk4(0.25, dt, 2)
        out.append((r, e.get_state()))
        e.close()
    assert np.array_equal(out[0][0], out[1][0])
    err = rel_err(out[0][1], out[1][1])
    assert np.all(err < TOL_PLANE[prec]), err


def test_bg_plane_off_for_k_varying_background():
    # a background that varies along k is read from the full arrays, and the
    # result still matches the oracle
    name, setup, kz = CASES[1]
    print(f"[case] k-varying background {name}", flush=True)
    bg = setup.bg.copy()
    n0, n1, n2 = setup.n
    kk = np.arange(n2).reshape(n2, 1, 1)
    fac = (1.0 + 1e-3 * np.sin(kk)).repeat(n1, 1).repeat(n0, 2).reshape(-1)
    bg[1] *= fac  # rho_b
    s = dataclasses.replace(setup, bg=bg)
    o = Oracle(s)
    w0 = o.boundary(perturbed(s, amp=1e-3, seed=10))
    e = Engine(s, FP64)
    e.set_zstream(True, kz)
    e.set_state(w0, bg)
    assert not e.get_bg_plane()
    got = e.rhs(t=0.5)
    e.close()
    ref = o.rhs(w0, 0.5)
    err = rel_err(got, ref, interior_mask(s))
    assert np.all(err < TOL_ORACLE_FP64), err


def _uniform_bg(s):
    """the set-up with its background replaced by the values of one interior
    cell everywhere and no gravity: a background that is periodic along
    every dim"""
    bg = np.repeat(s.bg[:, [s.n[0] * 3 + 2]], s.ncell, axis=1)
    return dataclasses.replace(s, bg=bg, g=(0.0, 0.0, 0.0))


@pytest.mark.parametrize("bc,uniform", [((4, 0, 0), False), ((4, 4, 0), False),
                                        ((4, 0, 4), False), ((0, 0, 0), True),
                                        ((0, 4, 4), True)],
                         ids=["cpp", "ccp", "cpc", "ppp_uniform", "pcc_uniform"])
def test_folded_ghost_fill_periodic_vs_oracle(bc, uniform):
    # the k-plane background path fills the ghosts of RK stage outputs in the
    # ghost-cell share of the folded aux pass (k_ghost_prep); with periodic
    # faces the images wrap around.  A periodic vertical axis needs a
    # background that is periodic along it (here uniform): the stratified
    # one is rejected (test_periodic_axis_needs_periodic_background)
    s = dataclasses.replace(strat(20, 12, 16), bc=bc)
    if uniform:
        s = _uniform_bg(s)
    print(f"[case] folded ghost fill bc={bc} n={s.n}", flush=True)
    o = Oracle(s)
    w0 = o.boundary(perturbed(s, amp=1e-3, seed=14))
    dt = 0.5 * configs.hyperviscous_dt(s)
    ref, _ = o.run(w0, 0.25, dt, 3, rk4=True)
    e = Engine(s, FP64)
    e.set_state(w0)
    assert e.get_bg_plane()
    e.rk4(0.25, dt, 3)
    got = e.get_state()
    e.close()
    # every cell, ghosts included
    err = rel_err(got, ref)
    assert np.all(err < TOL_ORACLE_FP64), err


def test_periodic_axis_needs_periodic_background():
    # a periodic vertical axis over the stratified background would give the
    # top ghost cells the bottom's density perturbation on the corona's
    # background (a negative full density): set_state rejects it
    s = dataclasses.replace(strat(20, 12, 16), bc=(0, 4, 4))
    print(f"[case] periodic vertical axis, stratified background n={s.n}", flush=True)
    e = Engine(s, FP64)
    with pytest.raises(SmaugError, match="dim 0 is periodic but background field"):
        e.set_state(s.w, s.bg)
    e.close()


@pytest.mark.parametrize("prec", [FP64, MIXED], ids=["fp64", "mixed"])
def test_steps_split_across_calls_match(prec):
    # with the k-plane background every stage's aux pass is done by the
    # update before it, and a step hands the aux quantities of its result to
    # the next: steps split over several calls, with calls between them that
    # recompute or replace the aux quantities, give the same bits as one call
    name, s, kz = CASES[1]
    print(f"[case] split rk4 calls {name} prec={prec}", flush=True)
    w0 = Oracle(s).boundary(perturbed(s, amp=1e-3, seed=16))
    dt = 0.5 * configs.hyperviscous_dt(s)
    e = Engine(s, prec)
    e.set_state(w0)
    assert e.get_bg_plane()
    e.rk4(0.0, dt, 5)
    ref = e.get_state()
    e.set_state(w0)
    e.rk4(0.0, dt, 2)
    e.cmax()
    e.rk4(2 * dt, dt, 1)
    e.rhs(t=3 * dt)
    e.rk4(3 * dt, dt, 1)
    e.rk4(4 * dt, dt, 1)
    got = e.get_state()
    e.close()
    assert np.array_equal(got, ref), rel_err(got, ref)


# the launch shapes the timed runs use: update and source z-stream depths 32,
# 64 and 128, ratio depths 4, 6, 8, 16 and 32 (the autotuner's picks at 256^3
# and 768x512x512, mixed and FP64), on a thin grid deep enough along k for
# several chunks
THIN = strat(12, 12, 260)


@pytest.mark.parametrize("kz,kzr", [(64, 8), (128, 16), (128, 8), (64, 16), (128, 4), (32, 32),
                                     (128, 6)],
                         ids=["kz64_r8", "kz128_r16", "kz128_r8", "kz64_r16", "kz128_r4",
                              "kz32_r32", "kz128_r6"])
def test_timed_shapes_rk4_fp64_vs_oracle(kz, kzr):
    print(f"[case] timed shape kz={kz} kzr={kzr} n={THIN.n}", flush=True)
    o = Oracle(THIN)
    w0 = o.boundary(perturbed(THIN, amp=1e-3, seed=17))
    dt = 0.5 * configs.hyperviscous_dt(THIN)
    ref, _ = o.run(w0, 0.25, dt, 3, rk4=True)
    e = Engine(THIN, FP64)
    e.set_zstream(True, kz)
    e.set_ratio_form(True, kzr)
    assert e.get_zstream() == (1, kz) and e.get_ratio_form() == (1, kzr)
    e.set_state(w0)
    e.rk4(0.25, dt, 3)
    got = e.get_state()
    e.close()
    err = rel_err(got, ref)
    print(f"kz={kz} kzr={kzr} rk4 rel err vs oracle", err)
    assert np.all(err < TOL_ORACLE_FP64), err


@pytest.mark.parametrize("kz,kzr", [(64, 8), (128, 8), (128, 4), (32, 8), (128, 6)],
                         ids=["kz64_r8", "kz128_r8", "kz128_r4", "kz32_r8", "kz128_r6"])
def test_timed_shapes_mixed_matches_per_cell(kz, kzr):
    print(f"[case] timed shape mixed kz={kz} kzr={kzr} n={THIN.n}", flush=True)
    o = Oracle(THIN)
    w = o.boundary(perturbed(THIN, amp=1e-2, seed=18))
    w = o.boundary(w.astype(np.float32).astype(np.float64))
    e = Engine(THIN, MIXED)
    e.set_state(w)
    e.set_zstream(False)
    e.set_ratio_form(False)
    ref = e.rhs(t=2.0)
    e.set_zstream(True, kz)
    e.set_ratio_form(True, kzr)
    got = e.rhs(t=2.0)
    e.close()
    err = rel_err(got, ref, interior_mask(THIN))
    print(f"kz={kz} kzr={kzr} mixed rhs rel err vs per-cell", err)
    assert np.all(err < TOL_RHS[MIXED]), err


# without hyperdiffusion the engine allocates no source buffer; the
# z-streamed update, whose lanes issue their source loads on every path,
# then reads a stand-in buffer (a grid with hyperdif = 0 once faulted here)
@pytest.mark.parametrize("prec", [FP64, MIXED], ids=["fp64", "mixed"])
def test_zstream_without_hyperdiffusion(prec):
    s = dataclasses.replace(strat(20, 12, 16), hyperdif=0)
    print(f"[case] zstream without hyperdiffusion prec={prec} n={s.n}", flush=True)
    o = Oracle(s)
    w0 = o.boundary(perturbed(s, amp=1e-3, seed=41))
    if prec == MIXED:
        w0 = o.boundary(w0.astype(np.float32).astype(np.float64))
    dt = 0.5 * configs.hyperviscous_dt(s)
    out = {}
    for form in ("zstream", "per-cell"):
        e = Engine(s, prec)
        if form == "per-cell":
            e.set_zstream(False)
        else:
            e.set_zstream(True, 5)
            assert e.get_zstream() == (1, 5)
        e.set_state(w0)
        e.rk4(0.5, dt, 3)
        out[form] = e.get_state()
        e.close()
    assert np.all(np.isfinite(out["zstream"]))
    if prec == FP64:
        ref, _ = o.run(w0, 0.5, dt, 3, rk4=True)
        err = rel_err(out["zstream"], ref)
        print("no-hyperdiffusion zstream rk4 rel err vs oracle", err)
        assert np.all(err < TOL_ORACLE_FP64), err
    err = rel_err(out["zstream"], out["per-cell"], interior_mask(s))
    print("no-hyperdiffusion zstream vs per-cell rel err", err)
    assert np.all(err < TOL_PLANE[prec] * 10), err
