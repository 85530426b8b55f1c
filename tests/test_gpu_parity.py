"""Parity of the HIP path (through the C-ABI of smaug_engine.h) with the CPU
oracle on identical seeded inputs."""
import numpy as np
import pytest

from helpers import interior_mask, perturbed, rel_err, rel_l2

pytestmark = pytest.mark.gpu

from smaugamd import configs  # noqa: E402
from smaugamd.engine import FP64, MIXED, Engine  # noqa: E402
from oracle.oracle import Oracle  # noqa: E402

# FP64 path vs FP64 oracle: same arithmetic up to operation order and FMA
# contraction, so agreement to a few ulps times the operator's condition
TOL_RHS_FP64 = 1e-10
# mixed path (FP32 perturbations) vs FP64 oracle, one evaluation of L(w)
TOL_RHS_MIXED = 2e-4


def small_setups():
    out = []
    out.append(("ot32", configs.orszag_tang(nx=32)))
    out.append(("bw64", configs.brio_wu(ncells=64)))
    s = configs.stratified(nz=20, nx=12, ny=16, lz=6e6 * 20 / 256 * 4, lx=4e6 * 12 / 256,
                           ly=4e6 * 16 / 256)
    out.append(("strat20x12x16", s))
    return out


@pytest.mark.parametrize("name,setup", small_setups(), ids=lambda x: x if isinstance(x, str) else "")
def test_rhs_fp64(name, setup):
    print(f"[case] rhs fp64 {name} n={setup.n}", flush=True)
    o = Oracle(setup)
    w = o.boundary(perturbed(setup, seed=1))
    ref = o.rhs(w, t=3.0)
    e = Engine(setup, FP64)
    e.set_state(w)
    got = e.rhs(t=3.0)
    err = rel_err(got, ref, interior_mask(setup))
    print(name, "rhs rel err per field", err)
    assert np.all(err < TOL_RHS_FP64), err


@pytest.mark.parametrize("name,setup", small_setups(), ids=lambda x: x if isinstance(x, str) else "")
def test_cmax_fp64(name, setup):
    print(f"[case] cmax {name}", flush=True)
    o = Oracle(setup)
    w = o.boundary(perturbed(setup, seed=2))
    e = Engine(setup, FP64)
    e.set_state(w)
    np.testing.assert_allclose(e.cmax()[:setup.ndim], o.cmax(w)[:setup.ndim], rtol=1e-13)


@pytest.mark.parametrize("name,setup", small_setups(), ids=lambda x: x if isinstance(x, str) else "")
def test_rhs_mixed(name, setup):
    print(f"[case] rhs mixed {name}", flush=True)
    o = Oracle(setup)
    w = o.boundary(perturbed(setup, seed=3))
    # the mixed path stores the perturbations in FP32: feed the oracle the
    # same rounded state so the comparison measures the arithmetic only
    w32 = o.boundary(w.astype(np.float32).astype(np.float64))
    ref = o.rhs(w32, t=1.0)
    e = Engine(setup, MIXED)
    e.set_state(w32)
    got = e.rhs(t=1.0)
    err = rel_err(got, ref, interior_mask(setup))
    print(name, "mixed rhs rel err per field", err)
    assert np.all(err < TOL_RHS_MIXED), err


@pytest.mark.parametrize("name,setup", small_setups(), ids=lambda x: x if isinstance(x, str) else "")
def test_rk4_steps_fp64(name, setup):
    print(f"[case] rk4 fp64 {name}", flush=True)
    o = Oracle(setup)
    w0 = o.boundary(perturbed(setup, amp=1e-3, seed=4))
    dt = 0.3 / configs.initial_max_courant(setup) if setup.dt == 0 else setup.dt * 0.5
    if name.startswith("ot"):
        dt = 1e-3
    ref, _ = o.run(w0, 0.0, dt, 5, rk4=True)
    e = Engine(setup, FP64)
    e.set_state(w0)
    e.rk4(0.0, dt, 5)
    got = e.get_state()
    err = rel_err(got, ref)
    print(name, "rk4 5 steps rel err", err)
    assert np.all(err < 1e-11), err



# mixed path after five RK4 steps against the FP64 oracle run from the same
# FP32-rounded state, per field, normalised by the field's amplitude: a stated
# bound on the FP32 stencil arithmetic and state storage over several steps
# (measured: at most 7e-7 L-infinity and 4.5e-7 L2 on these grids)
TOL_RK4_MIXED_LINF = 1e-5
TOL_RK4_MIXED_L2 = 5e-6


@pytest.mark.parametrize("name,setup", small_setups(), ids=lambda x: x if isinstance(x, str) else "")
def test_rk4_steps_mixed_vs_oracle(name, setup):
    print(f"[case] rk4 mixed vs oracle {name}", flush=True)
    o = Oracle(setup)
    w0 = o.boundary(perturbed(setup, amp=1e-3, seed=4))
    w0 = o.boundary(w0.astype(np.float32).astype(np.float64))
    dt = 0.3 / configs.initial_max_courant(setup) if setup.dt == 0 else setup.dt * 0.5
    if name.startswith("ot"):
        dt = 1e-3
    ref, _ = o.run(w0, 0.0, dt, 5, rk4=True)
    e = Engine(setup, MIXED)
    e.set_state(w0)
    e.rk4(0.0, dt, 5)
    got = e.get_state()
    e.close()
    m = interior_mask(setup)
    linf, l2 = rel_err(got, ref, m), rel_l2(got, ref, m)
    print(name, "mixed rk4 5 steps rel Linf", linf, "rel L2", l2)
    assert np.all(np.isfinite(got))
    assert np.all(linf < TOL_RK4_MIXED_LINF), linf
    assert np.all(l2 < TOL_RK4_MIXED_L2), l2

def test_euler_steps_fp64():
    setup = configs.brio_wu(ncells=64)
    print("[case] euler bw64", flush=True)
    o = Oracle(setup)
    w0 = o.boundary(setup.w)
    ref, _ = o.run(w0, 0.0, 2e-4, 10, rk4=False)
    e = Engine(setup, FP64)
    e.set_state(w0)
    e.euler(0.0, 2e-4, 10)
    got = e.get_state()
    err = rel_err(got, ref)
    print("euler 10 steps rel err", err)
    assert np.all(err < 1e-11), err


@pytest.mark.parametrize("form", ["zstream", "per-cell"])
def test_torsional_driver_fp64(form):
    # the reference's own user driver (usersource.cu), both 3-D update forms
    s = configs.torsional_driver(configs.stratified(nz=20, nx=12, ny=16, lz=6e6 * 20 / 256 * 4,
                                                    lx=4e6 * 12 / 256, ly=4e6 * 16 / 256))
    print(f"[case] torsional driver fp64 form={form} n={s.n}", flush=True)
    o = Oracle(s)
    w = o.boundary(perturbed(s, amp=1e-3, seed=13))
    t = 6.0  # sin(2 pi t / 30) != 0
    ref_r = o.rhs(w, t=t)
    ref, _ = o.run(w, t, s.dt, 3, rk4=True)
    e = Engine(s, FP64)
    if form == "per-cell":
        e.set_zstream(False)
    e.set_state(w)
    got_r = e.rhs(t=t)
    e.rk4(t, s.dt, 3)
    got = e.get_state()
    e.close()
    m = interior_mask(s)
    # the driver moves the horizontal momenta
    assert np.abs(ref_r[2][m]).max() > 0 and np.abs(ref_r[3][m]).max() > 0
    err = rel_err(got_r, ref_r, m)
    assert np.all(err < TOL_RHS_FP64), err
    err = rel_err(got, ref, m)
    assert np.all(err < TOL_RHS_FP64), err


def test_torsional_driver_rejected_in_2d():
    s = configs.orszag_tang(nx=16)
    s.driver = 2
    s.drv = np.array([500.0, 30.0, 0.1, 0.0, 0.1, 0.5, 0.5, 4.0])
    with pytest.raises(Exception, match="3-D only"):
        Engine(s, FP64)


@pytest.mark.parametrize("name", ["ot32-stretched", "strat-stretched"])
def test_maxcourant_fp64_nonuniform(name):
    # max over cells of cfast_d / dx_d at the cell: on a stretched grid this
    # differs from max cfast_d / min dx_d
    if name.startswith("ot"):
        s = configs.orszag_tang(nx=32)
    else:
        s = configs.stratified(nz=20, nx=12, ny=16, lz=6e6 * 20 / 256 * 4, lx=4e6 * 12 / 256,
                               ly=4e6 * 16 / 256)
    d = s.ndim - 1
    x = np.arange(s.n[d])
    s.dx[d] = s.dx[d] * (1.0 + 0.3 * np.sin(2 * np.pi * x / s.n[d]))
    print(f"[case] maxcourant stretched {name} n={s.n}", flush=True)
    o = Oracle(s)
    w = o.boundary(perturbed(s, amp=1e-2, seed=15))
    ref = o.maxcourant(w)
    e = Engine(s, FP64)
    e.set_state(w)
    got = e.maxcourant()
    cm = e.cmax()
    e.close()
    uniform_bound = max(cm[k] / np.min(s.dx[k]) for k in range(s.ndim))
    assert got == pytest.approx(ref, rel=1e-12)
    assert got < uniform_bound * (1 - 1e-6)  # the stretched axis matters


def test_mixed_rk4_200_steps_vs_oracle_stratified():
    """the mixed path after 200 RK4 steps of a reduced stratified workload
    (128 x 16 x 16 interior cells, the workload's cell sizes and dt, the
    transition region included, driver on) against the FP64 oracle: relative
    L2 per field.  Measured on MI355X: rho 3.5e-7, m_z 1.8e-7, m_x and m_y
    1.7e-6, e 4.0e-7, b 1.5e-6 to 1.6e-6; the bounds below are about three
    times those."""
    full = configs.stratified(nz=256, nx=8, ny=8, lx=4e6 * 8 / 256, ly=4e6 * 8 / 256)
    s = configs.stratified(nz=128, nx=16, ny=16, lz=6e6 * 128 / 256, lx=4e6 * 16 / 256,
                           ly=4e6 * 16 / 256)
    s.dt = full.dt
    n = 200
    print(f"[case] mixed vs oracle, {n} RK4 steps, n={s.n}, dt={s.dt:.4e}", flush=True)
    o = Oracle(s)
    w0 = o.boundary(s.w)
    ref, _ = o.run(w0, 0.0, s.dt, n, rk4=True)
    e = Engine(s, MIXED)
    e.set_state(w0)
    e.rk4(0.0, s.dt, n)
    got = e.get_state()
    e.close()
    m = interior_mask(s)
    l2 = rel_l2(got, ref, m)
    names = ["rho", "m1", "m2", "m3", "e", "b1", "b2", "b3"]
    print("mixed vs oracle after 200 steps, rel L2:",
          {k: f"{v:.2e}" for k, v in zip(names, l2)}, flush=True)
    assert np.all(np.isfinite(got))
    bound = np.array([1.5e-6, 1.0e-6, 6e-6, 6e-6, 1.5e-6, 6e-6, 6e-6, 6e-6])
    assert np.all(l2 < bound), l2


@pytest.fixture(scope="module")
def strat64_100_steps():
    """a 64 x 32 x 32 stratified grid (the workload's cell sizes and dt,
    driver on, a seeded perturbation of every field) and the FP64 oracle's
    state after 100 RK4 steps"""
    full = configs.stratified(nz=256, nx=8, ny=8, lx=4e6 * 8 / 256, ly=4e6 * 8 / 256)
    s = configs.stratified(nz=64, nx=32, ny=32, lz=6e6 * 64 / 256, lx=4e6 * 32 / 256,
                           ly=4e6 * 32 / 256)
    s.dt = full.dt
    o = Oracle(s)
    w0 = o.boundary(perturbed(s, amp=1e-3, seed=41))
    ref, _ = o.run(w0, 0.0, s.dt, 100, rk4=True)
    return s, w0, ref


@pytest.mark.parametrize("shape", ["default", "kz16_r8"])
def test_fp64_rk4_100_steps_vs_oracle_stratified(strat64_100_steps, shape):
    """the FP64 path after 100 RK4 steps against the FP64 oracle, in the
    default launch shapes (the in-plane ratios formed by the source kernel)
    and with two z-chunks for the update and sources and ratio chunks of 8
    planes.  Measured on MI355X: 3.3e-15 of each field's amplitude at most
    (L-infinity, every cell); the bound is 1e-12."""
    s, w0, ref = strat64_100_steps
    print(f"[case] fp64 vs oracle, 100 RK4 steps, n={s.n}, shape={shape}", flush=True)
    e = Engine(s, FP64)
    if shape != "default":
        e.set_zstream(1, 16)
        e.set_ratio_form(2, 8)
    assert e.get_ratio_form()[0] == 2
    e.set_state(w0)
    e.rk4(0.0, s.dt, 100)
    got = e.get_state()
    e.close()
    err = rel_err(got, ref)
    print("fp64 vs oracle after 100 steps, rel Linf per field:", err, flush=True)
    assert np.all(err < 1e-12), err


@pytest.fixture(scope="module")
def ot128_runs():
    """the Orszag-Tang vortex on 128 x 128 cells (periodic, dt = 1e-3): the
    FP64 oracle's states after 20 and 200 RK4 steps, and its own spread after
    200 steps -- the change a seeded 1e-15 relative perturbation of the
    initial state makes.  The vortex amplifies round-off about 1e9-fold over
    200 steps (measured spread: 8.8e-13 at 20 steps, 3.4e-6 at 200)."""
    s = configs.orszag_tang(nx=128)
    o = Oracle(s)
    w0 = o.boundary(s.w)
    ref20, _ = o.run(w0, 0.0, 1e-3, 20, rk4=True)
    ref200, _ = o.run(w0, 0.0, 1e-3, 200, rk4=True)
    rng = np.random.default_rng(1)
    w1 = o.boundary(w0 * (1 + 1e-15 * rng.standard_normal(w0.shape)))
    alt200, _ = o.run(w1, 0.0, 1e-3, 200, rk4=True)
    return s, w0, ref20, ref200, rel_err(alt200, ref200)


@pytest.mark.parametrize("prec", [FP64, MIXED], ids=["fp64", "mixed"])
def test_rk4_long_vs_oracle_orszag_tang(ot128_runs, prec):
    """the 2-D per-cell path after 20 and 200 RK4 steps of the Orszag-Tang
    vortex against the FP64 oracle, per field, normalised by the field's
    amplitude (L-infinity, every cell).  FP64: to 1e-11 after 20 steps, and
    within the oracle's own round-off spread after 200.  Mixed: the same
    amplification acts on FP32's rounding (6e-8 per stored value); measured
    on MI355X 2.3e-6 after 20 steps and 4.1e-4 after 200, bounded at 1e-5
    and 2e-3."""
    s, w0, ref20, ref200, spread = ot128_runs
    print(f"[case] orszag-tang 128^2, 20 and 200 RK4 steps, prec={prec}", flush=True)
    e = Engine(s, prec)
    e.set_state(w0)
    e.rk4(0.0, 1e-3, 20)
    got20 = e.get_state()
    e.rk4(20e-3, 1e-3, 180)
    got200 = e.get_state()
    e.close()
    err20, err200 = rel_err(got20, ref20), rel_err(got200, ref200)
    print("orszag-tang rel Linf per field after 20 steps:", err20, flush=True)
    print("after 200 steps:", err200, "oracle spread:", spread, flush=True)
    assert np.all(np.isfinite(got200))
    if prec == FP64:
        assert np.all(err20 < 1e-11), err20
        assert np.all(err200 <= spread.max()), (err200, spread)
    else:
        assert np.all(err20 < 1e-5), err20
        assert np.all(err200 < 2e-3), err200
