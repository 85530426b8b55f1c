"""moddton through the drop-in path: the C driver (smaug-all_amd/host/
smaug_main.c) sets every onestep dt from the Courant limit
(cucomputemaxcourant) and, from the second step on, caps it by the
hyperviscous limit (cugetdtvisc1, from the p->maxviscoef that the previous
step's cuhyperdifvisc1r/1l calls left), as iosmaug.cpp:855-898 does.  The
oracle restatement of that loop is tests/moddton_ref.py.

The hyperviscous cap is well conditioned only where every window of every
ratio holds real gradients: the ratio max|D3|/max|D1| of a region that is
flat up to rounding (values of 1e-15 ahead of a Brio-Wu wave, say) is a
ratio of rounding errors, and when the largest coefficient of the grid sits
there, two correct FP64 evaluations give caps that differ in the sixth
digit.  The smooth and noise-perturbed states are therefore compared to
round-off, step by step; Brio-Wu to round-off over its first steps, whose
largest coefficients sit on the initial jump.  The last case is the Brio-Wu
workload itself: 1024 cells, CFL 0.3, run to t = 0.1 (the driver's -tend,
SAC's tmaxexact), about 1900 steps; its own docstring says how it is
compared."""
import re

import numpy as np
import pytest

from helpers import interior_mask, perturbed, rel_err, rel_l2
from moddton_ref import onestep_moddton
from oracle.oracle import Oracle
from smaugamd import configs, vacio
from test_gpu_dropin import _drv_arg, roundtrip_state, run_driver

pytestmark = pytest.mark.gpu

STEP_LINE = re.compile(r"smaug_main: (\d+) steps, t = \S+, dt = (\S+), elapsed = (\S+)")


def _parse(stdout):
    m = STEP_LINE.search(stdout)
    assert m, stdout
    return int(m.group(1)), float(m.group(2)), float(m.group(3))


def test_moddton_orszag_tang_32(tmp_path):
    s = configs.orszag_tang(nx=32)
    o = Oracle(s)
    s.w = o.boundary(s.w)
    inp = roundtrip_state(s, tmp_path, "ot_dt")
    out = str(tmp_path / "ot_dt_out.ini")
    dt0, n = 5e-3, 20
    stdout = run_driver(2, ["-i", inp, "-o", out, "-n", str(n), "-dt", repr(dt0), "-bc", "0,0",
                            "-chyp", "0.045,0.02,0.4,0.02", "-moddton", "-courant", "0.3"],
                        tmp_path)
    steps, dt_last, elapsed = _parse(stdout)
    _, _, data = vacio.read_ascii(out)
    ref, _, el_ref, dts, _, _ = onestep_moddton(s, s.w, dt0, courant=0.3, nmax=n)
    err = rel_err(data[: s.nve], ref)
    print(f"OT-32 moddton: dt {dts[0]:.9e} .. {dts[-1]:.9e}; driver last {dt_last:.9e}; "
          f"rel err {err}")
    assert steps == n
    assert abs(dt_last - dts[-1]) <= 1e-11 * dts[-1]
    assert abs(elapsed - el_ref) <= 1e-11 * el_ref
    assert np.all(err < 1e-10), err


def test_moddton_brio_wu_64(tmp_path):
    s = configs.brio_wu(ncells=64)
    inp = roundtrip_state(s, tmp_path, "bw64")
    out = str(tmp_path / "bw64_out.ini")
    dt0, n = 1e-3, 4
    stdout = run_driver(2, ["-i", inp, "-o", out, "-n", str(n), "-dt", repr(dt0), "-bc", "4,4",
                            "-moddton", "-courant", "0.3"], tmp_path)
    steps, dt_last, elapsed = _parse(stdout)
    _, _, data = vacio.read_ascii(out)
    ref, nref, el_ref, dts, _, _ = onestep_moddton(s, s.w, dt0, courant=0.3, nmax=n)
    err = rel_err(data[: s.nve], ref)
    print(f"BW-64 moddton: dt {dts}; driver last {dt_last!r}; rel err {err}")
    assert steps == nref == n
    assert abs(dt_last - dts[-1]) <= 1e-12 * dts[-1]
    assert abs(elapsed - el_ref) <= 1e-12 * el_ref
    # the hyperviscous cap acts from the third step on, the Courant dt before it
    assert dts[2] < 0.3 / Oracle(s).maxcourant_dim(ref, 1)
    assert np.all(err < 1e-11), err


def test_moddton_stratified_small(tmp_path):
    s = configs.stratified(nz=20, nx=12, ny=12, lz=6e6 * 20 / 256 * 4, lx=4e6 * 12 / 256,
                           ly=4e6 * 12 / 256)
    o = Oracle(s)
    s.w = o.boundary(perturbed(s, amp=1e-3, seed=21))
    inp = roundtrip_state(s, tmp_path, "strat_dt")
    out = str(tmp_path / "strat_dt_out.ini")
    n, t0 = 4, 3.0
    stdout = run_driver(3, ["-i", inp, "-o", out, "-n", str(n), "-dt", repr(s.dt), "-t0", str(t0),
                            "-bc", "4,4,4", "-driver", _drv_arg(s), "-moddton"], tmp_path)
    steps, dt_last, elapsed = _parse(stdout)
    _, _, data = vacio.read_ascii(out)
    ref, _, el_ref, dts, _, _ = onestep_moddton(s, s.w, s.dt, courant=0.3, nmax=n, t0=t0)
    err = rel_err(data[: s.nve], ref)
    print(f"stratified moddton dts {dts}; driver last {dt_last!r}; rel err {err}")
    assert steps == n
    assert abs(dt_last - dts[-1]) <= 1e-11 * dts[-1]
    assert np.all(err < 1e-10), err


def test_brio_wu_1024_cfl03_to_t01(tmp_path):
    """BASELINE configuration 1 on the GPU drop-in path: Brio-Wu, 1024 cells,
    gamma 2, CFL 0.3 with the hyperviscous cap, forward Euler (the rkon = 0
    loop) to t = 0.1.  The initial dt (which the reference keeps for the
    energy field's own hyperdiffusion) is 8e-5, about the first Courant step.

    This run is chaotic at the round-off level in its second half: with the
    same dt sequence, the oracle started from a state perturbed by 1e-15
    differs from its unperturbed run by about 3e-12 after 800 steps and by
    about 1e-3 at t = 0.1 (about 1900 steps).  So:
      - over the first 800 steps the driver's state is compared with the
        oracle replaying the driver's logged dt sequence, to round-off, and
        each logged dt with the limits the oracle evaluates on the replayed
        state: the Courant dt to round-off where it is the smaller by a
        margin; where the hyperviscous cap acts, it is ill-conditioned (the
        grid's largest coefficient sits ahead of the waves, on values of
        1e-15, a ratio of rounding errors), so the dt is only held between
        the cap's floor and the Courant dt.  The floor: |D3| =
        |2 D1(f) - D1(f+1) - D1(f-1)| <= 4 max|D1| over the same window, so
        every coefficient is at most 4 dx cmax_d chyp and the cap at least
        0.25 min dx^2 / (4 max_d(dx cmax_d) max chyp);
      - at t = 0.1 the driver's deviation from the replay must stay within
        a few times the oracle's own spread, measured here;
      - the solution keeps its invariants."""
    s = configs.brio_wu(ncells=1024)
    inp = roundtrip_state(s, tmp_path, "bw1024")
    dt0 = 8.0e-5
    common = ["-i", inp, "-dt", repr(dt0), "-bc", "4,4", "-moddton", "-courant", "0.3"]
    out = str(tmp_path / "bw1024_out.ini")
    log = str(tmp_path / "bw1024_dt.txt")
    steps, _, elapsed = _parse(run_driver(2, common + ["-o", out, "-tend", "0.1", "-dtlog", log],
                                          tmp_path))
    _, _, data = vacio.read_ascii(out)
    logged = np.loadtxt(log)
    assert logged.shape == (steps, 2) and np.all(logged[:, 0] == np.arange(1, steps + 1))
    gdt = logged[:, 1]
    assert abs(elapsed - 0.1) <= 1e-15 and abs(gdt.sum() - 0.1) <= 1e-14
    # the first 800 steps
    K = 800
    out_k = str(tmp_path / "bw1024_out_k.ini")
    run_driver(2, common + ["-o", out_k, "-n", str(K)], tmp_path)
    _, _, data_k = vacio.read_ascii(out_k)
    ref_k, _, _, _, cour, visc = onestep_moddton(s, s.w, dt0, courant=0.3, dts_given=gdt[:K])
    err_k = rel_err(data_k[: s.nve], ref_k)
    capped = (visc > 1e-8) & (visc < cour)
    expect = np.where(capped, visc, cour)
    dev = np.abs(gdt[:K] - expect) / expect
    # Courant-limited by a margin: an ill-conditioned cap cannot have
    # switched the limit there
    tight = ~((visc > 1e-8) & (visc < cour * (1 + 1e-3)))
    # the cap's floor on every replayed state (cmax of the state before the step)
    o = Oracle(s)
    w = o.boundary(s.w)
    hmin = min(float(dx.min()) for dx in s.dx)
    floor = np.empty(K)
    for n in range(K):
        cm = o.cmax(w)
        floor[n] = 0.25 * hmin ** 2 / (4 * max(float(s.dx[d].max()) * cm[d] for d in range(2)) *
                                       float(s.chyp.max()))
        w, *_ = onestep_moddton(s, w, dt0, courant=0.3, dts_given=gdt[n:n + 1])
    # t = 0.1: the replay of the whole sequence, and the oracle's own spread
    ref, *_ = onestep_moddton(s, s.w, dt0, courant=0.3, dts_given=gdt)
    rng = np.random.default_rng(3)
    w_p = s.w * (1 + 1e-15 * rng.standard_normal(s.w.shape))
    ref_p, *_ = onestep_moddton(s, w_p, dt0, courant=0.3, dts_given=gdt)
    err = rel_err(data[: s.nve], ref).max()
    spread = rel_err(ref_p, ref).max()
    q = np.quantile(dev[capped], [0.5, 0.9, 1.0]) if capped.any() else [0, 0, 0]
    print(f"BW-1024 to t=0.1: {steps} steps; first {K}: rel err {err_k.max():.2e}, dt dev "
          f"Courant-limited {dev[tight].max() if tight.any() else 0:.2e} ({tight.sum()} steps), "
          f"capped median/90%/max {q[0]:.1e}/{q[1]:.1e}/{q[2]:.1e} ({capped.sum()} steps); "
          f"t=0.1: rel err {err:.2e}, oracle spread {spread:.2e}")
    assert np.all(err_k < 1e-9), err_k
    assert not tight.any() or dev[tight].max() < 1e-11
    assert np.all(gdt[:K] <= cour * (1 + 1e-9))
    assert np.all(gdt[:K] >= floor * (1 - 1e-9))
    assert err <= 5 * spread + 1e-9
    # the solution itself: density between the two initial states up to the
    # hyperdiffusion's small undershoot, mass conserved (the waves stay
    # inside the tube by t = 0.1)
    n0 = s.n[0]
    rho = data[0].reshape(s.n[1], n0)[4, 2:-2]
    assert 0.11 < rho.min() and rho.max() < 1.0 + 1e-6
    rho0 = s.w[0].reshape(s.n[1], n0)[4, 2:-2]
    assert abs(rho.sum() - rho0.sum()) <= 1e-9 * rho0.sum()
