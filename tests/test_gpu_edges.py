"""Edge sizes of the grid: the smallest grid the engine accepts (8 cells per
dimension including the two ghost layers on each side, so 4 interior cells),
grids thinner than one 64 x 4 tile along one axis and ragged in the others,
in every form of the 3-D stage update and of the ratio pass, against the
oracle."""
import numpy as np
import pytest

from helpers import interior_mask, perturbed, rel_err

pytestmark = pytest.mark.gpu

from smaugamd import configs  # noqa: E402
from smaugamd.engine import FP64, MIXED, Engine  # noqa: E402
from oracle.oracle import Oracle  # noqa: E402

TOL_FP64 = 1e-10
# one evaluation of L(w) on FP32 perturbations against the FP64 oracle
TOL_RHS_MIXED = 2e-4


def strat(nz, nx, ny):
    return configs.stratified(nz=nz, nx=nx, ny=ny, lz=6e6 * nz / 256 * 4,
                              lx=4e6 * nx / 256 * 4, ly=4e6 * ny / 256 * 4)


SHAPES_3D = [(4, 4, 4), (4, 65, 5), (67, 4, 6)]
# (update form, depth): per cell, and z-streamed with each source kernel at
# the default depth and at a depth that does not divide 4 planes
FORMS = [(0, 0), (1, 0), (1, 3), (2, 3), (3, 0)]


@pytest.mark.parametrize("shape", SHAPES_3D, ids=lambda s: "x".join(map(str, s)))
def test_thin_3d_grids_rk4_vs_oracle(shape):
    s = strat(*shape)
    print(f"[case] thin 3-D grid rk4 n={s.n}", flush=True)
    o = Oracle(s)
    w0 = o.boundary(perturbed(s, amp=1e-3, seed=21))
    dt = 0.5 * configs.hyperviscous_dt(s)
    ref, _ = o.run(w0, 0.25, dt, 3, rk4=True)
    m = interior_mask(s)
    for mode, kz in FORMS:
        for ratio_z in (2, 1, 0):
            print(f"[case] thin 3-D grid n={s.n} form={mode} kz={kz} "
                  f"ratios z-streamed={ratio_z}", flush=True)
            e = Engine(s, FP64)
            e.set_zstream(mode, kz)
            e.set_ratio_form(ratio_z, kz)
            e.set_state(w0)
            e.rk4(0.25, dt, 3)
            got = e.get_state()
            e.close()
            err = rel_err(got, ref, m)
            assert np.all(np.isfinite(got))
            assert np.all(err < TOL_FP64), (mode, kz, ratio_z, err)


@pytest.mark.parametrize("shape", SHAPES_3D, ids=lambda s: "x".join(map(str, s)))
def test_thin_3d_grids_mixed_rhs(shape):
    s = strat(*shape)
    print(f"[case] thin 3-D grid mixed rhs n={s.n}", flush=True)
    o = Oracle(s)
    w = o.boundary(perturbed(s, amp=1e-2, seed=22))
    w32 = o.boundary(w.astype(np.float32).astype(np.float64))
    ref = o.rhs(w32, t=1.5)
    e = Engine(s, MIXED)
    e.set_state(w32)
    got = e.rhs(t=1.5)
    e.close()
    err = rel_err(got, ref, interior_mask(s))
    assert np.all(err < TOL_RHS_MIXED), err


@pytest.mark.parametrize("prec", [FP64, MIXED], ids=["fp64", "mixed"])
def test_smallest_2d_grid_rk4_vs_oracle(prec):
    s = configs.orszag_tang(nx=4)
    print(f"[case] smallest 2-D grid prec={prec} n={s.n}", flush=True)
    o = Oracle(s)
    w0 = o.boundary(perturbed(s, amp=1e-2, seed=23))
    if prec == MIXED:
        w0 = o.boundary(w0.astype(np.float32).astype(np.float64))
    dt = 0.2 / configs.initial_max_courant(s)
    ref, _ = o.run(w0, 0.0, dt, 4, rk4=True)
    e = Engine(s, prec)
    e.set_state(w0)
    e.rk4(0.0, dt, 4)
    got = e.get_state()
    e.close()
    err = rel_err(got, ref, interior_mask(s))
    # four steps of FP32 arithmetic against FP64
    assert np.all(err < (TOL_FP64 if prec == FP64 else 1e-4)), err

