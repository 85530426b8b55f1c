"""The 3-D fast path at the full size of the stratified workload (256^3
interior cells), through properties that do not depend on the grid size:
the z-streamed kernels (k-plane background, ghost fill folded into the aux
pass) against the per-cell kernels (full background, separate ghost fill),
which share no stage code beyond the pointwise physics, for L(w) and for
RK4 steps."""
import numpy as np
import pytest

from helpers import interior_mask, perturbed, rel_err

pytestmark = pytest.mark.gpu

from smaugamd import configs  # noqa: E402
from smaugamd.engine import FP64, MIXED, Engine  # noqa: E402
from oracle.oracle import Oracle  # noqa: E402

# the two forms differ in the summation order of the neighbour contributions
TOL_RHS = {FP64: 1e-12, MIXED: 2e-5}
TOL_RK4 = {FP64: 1e-12, MIXED: 1e-4}


@pytest.fixture(scope="module")
def full():
    s = configs.stratified(nz=256, nx=256, ny=256)
    o = Oracle(s)
    w = o.boundary(perturbed(s, amp=1e-3, seed=31))
    return s, w


@pytest.mark.parametrize("prec", [FP64, MIXED], ids=["fp64", "mixed"])
def test_fullsize_zstream_matches_per_cell(full, prec):
    s, w = full
    print(f"[case] full-size 256^3 z-streamed vs per-cell prec={prec}", flush=True)
    if prec == MIXED:
        w = w.astype(np.float32).astype(np.float64)
    m = interior_mask(s)
    out = {}
    for form in ("zstream", "per-cell"):
        e = Engine(s, prec)
        if form == "per-cell":
            e.set_zstream(False)
            e.set_ratio_form(False)
            e.set_bg_plane(False)
        e.set_state(w)
        assert e.get_bg_plane() == (form == "zstream")
        r = e.rhs(t=1.0)
        e.rk4(1.0, s.dt, 2)
        out[form] = (r, e.get_state())
        e.close()
    err = rel_err(out["zstream"][0], out["per-cell"][0], m)
    print("rhs rel err", err, flush=True)
    assert np.all(err < TOL_RHS[prec]), err
    assert np.all(np.isfinite(out["zstream"][1]))
    # every cell, ghosts included: the folded ghost fill leaves the state
    # between steps complete
    err = rel_err(out["zstream"][1], out["per-cell"][1])
    print("rk4 rel err", err, flush=True)
    assert np.all(err < TOL_RK4[prec]), err


@pytest.mark.timeout(900)
def test_large_grid_zstream_matches_per_cell_mixed():
    """the large stratified atmosphere at its own size, 768 x 512 x 512
    interior cells over 12 x 8 x 8 Mm (201 M cells, 62 GB per engine in
    mixed precision), with the launch shape its timed runs use (update and
    sources at z-stream depth 64, the ratios along k at depth 8 and the
    in-plane ones formed by the source kernel): two mixed RK4 steps of the
    z-streamed kernels against the per-cell kernels"""
    s = configs.stratified(nz=768, nx=512, ny=512, lz=12.0e6, lx=8.0e6, ly=8.0e6)
    print(f"[case] large grid {s.n} mixed, z-streamed vs per-cell", flush=True)
    rng = np.random.default_rng(33)
    w = s.w.copy()
    rho = s.bg[1]
    # small perturbations everywhere, float32 noise scaled per cell
    w[0] += 1e-3 * rho * rng.standard_normal(s.ncell, dtype=np.float32)
    w[4] += 1e-3 * s.bg[0] * rng.standard_normal(s.ncell, dtype=np.float32)
    w = w.astype(np.float32).astype(np.float64)
    m = interior_mask(s)
    out = {}
    for form in ("zstream", "per-cell"):
        e = Engine(s, MIXED)
        if form == "per-cell":
            e.set_zstream(False)
            e.set_ratio_form(False)
            e.set_bg_plane(False)
        else:
            e.set_zstream(True, 64)
            e.set_ratio_form(2, 8)
        e.set_state(w)
        e.rk4(1.0, s.dt, 2)
        out[form] = e.get_state()
        e.close()
        print(f"[case] large grid {form} done", flush=True)
    assert np.all(np.isfinite(out["zstream"]))
    err = rel_err(out["zstream"], out["per-cell"], m)
    print("large grid rk4 rel err", err, flush=True)
    assert np.all(err < TOL_RK4[MIXED]), err
