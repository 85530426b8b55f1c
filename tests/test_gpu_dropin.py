"""The drop-in path: the C driver (smaug-all_amd/host/smaug_main.c) runs the
reference's onestep time loop through the cu* entry points of
libsmaug_hip2d/3d.so, reading and writing VAC ascii configuration files;
its result is compared with the oracle's forward-Euler (rkon=0) run, and its
RK4 branches, fused (cuadvancerk4) and through the per-entry-point calls
with cuadvance's slot algebra (-rk4calls), with the oracle's RK4."""
import os
import subprocess

import numpy as np
import pytest

from helpers import perturbed, rel_err
from oracle.oracle import Oracle
from smaugamd import configs, vacio

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "smaug-all_amd", "lib")


def run_driver(ndim, args, tmp_path):
    exe = os.path.join(LIB, f"smaug_main{ndim}d")
    assert os.path.exists(exe), f"{exe} not built"
    print("[case] driver", exe, " ".join(args), flush=True)
    r = subprocess.run([exe] + args, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr + r.stdout
    return r.stdout


def roundtrip_state(s, tmp_path, name):
    """write the set-up as a VAC ascii file and read back exactly what the
    driver will read"""
    path = str(tmp_path / f"{name}.ini")
    vacio.write_ascii(path, s)
    _, pos, data = vacio.read_ascii(path)
    s.w = data[: s.nve].copy()
    s.bg = data[s.nve:].copy()
    return path


def test_driver_onestep_ot(tmp_path):
    s = configs.orszag_tang(nx=32)
    s.w = perturbed(s, amp=1e-3, seed=7)
    o = Oracle(s)
    s.w = o.boundary(s.w)
    inp = roundtrip_state(s, tmp_path, "ot")
    out = str(tmp_path / "ot_out.ini")
    dt, n = 1e-3, 4
    run_driver(2, ["-i", inp, "-o", out, "-n", str(n), "-dt", str(dt), "-bc", "0,0",
                   "-chyp", "0.045,0.02,0.4,0.02"], tmp_path)
    _, _, data = vacio.read_ascii(out)
    ref, _ = Oracle(s).run(s.w, dt, dt, n, rk4=False)
    err = rel_err(data[: s.nve], ref)
    print("driver onestep OT rel err", err)
    assert np.all(err < 1e-11), err


def _strat_small():
    return configs.stratified(nz=20, nx=12, ny=12, lz=6e6 * 20 / 256 * 4, lx=4e6 * 12 / 256,
                              ly=4e6 * 12 / 256)


def _drv_arg(s):
    return ",".join(repr(float(v)) for v in s.drv[:7])


def test_driver_onestep_stratified(tmp_path):
    s = _strat_small()
    o = Oracle(s)
    s.w = o.boundary(perturbed(s, amp=1e-3, seed=8))
    inp = roundtrip_state(s, tmp_path, "strat")
    out = str(tmp_path / "strat_out.ini")
    n, t0 = 3, 2.0
    run_driver(3, ["-i", inp, "-o", out, "-n", str(n), "-dt", repr(s.dt), "-t0", str(t0),
                   "-bc", "4,4,4", "-driver", _drv_arg(s)], tmp_path)
    _, _, data = vacio.read_ascii(out)
    # the driver advances p->qt before each step (iosmaug.cpp:918)
    ref, _ = Oracle(s).run(s.w, t0 + s.dt, s.dt, n, rk4=False)
    err = rel_err(data[: s.nve], ref)
    print("driver onestep stratified rel err", err)
    assert np.all(err < 1e-11), err


def test_driver_rk4_fused_stratified(tmp_path):
    s = _strat_small()
    o = Oracle(s)
    s.w = o.boundary(perturbed(s, amp=1e-3, seed=9))
    inp = roundtrip_state(s, tmp_path, "strat_rk")
    out = str(tmp_path / "strat_rk_out.ini")
    n, t0 = 3, 1.0
    run_driver(3, ["-i", inp, "-o", out, "-n", str(n), "-dt", repr(s.dt), "-t0", str(t0),
                   "-bc", "4,4,4", "-driver", _drv_arg(s), "-rk4"], tmp_path)
    _, _, data = vacio.read_ascii(out)
    ref, _ = Oracle(s).run(s.w, t0, s.dt, n, rk4=True)
    err = rel_err(data[: s.nve], ref)
    print("driver rk4 stratified rel err", err)
    assert np.all(err < 1e-11), err


def test_driver_rk4calls_ot(tmp_path):
    # RK4 through the entry points: four stages of the onestep call sequence
    # (dt/2, dt/2, dt, dt/6) and cuadvance(order) after each
    s = configs.orszag_tang(nx=32)
    s.w = perturbed(s, amp=1e-3, seed=11)
    s.w = Oracle(s).boundary(s.w)
    inp = roundtrip_state(s, tmp_path, "ot_rkc")
    out = str(tmp_path / "ot_rkc_out.ini")
    dt, n = 1e-3, 4
    run_driver(2, ["-i", inp, "-o", out, "-n", str(n), "-dt", str(dt), "-bc", "0,0",
                   "-chyp", "0.045,0.02,0.4,0.02", "-rk4calls"], tmp_path)
    _, _, data = vacio.read_ascii(out)
    ref, _ = Oracle(s).run(s.w, 0.0, dt, n, rk4=True)
    err = rel_err(data[: s.nve], ref)
    print("driver rk4calls OT rel err", err)
    assert np.all(err < 1e-11), err
    # one Euler step differs from RK4 by O(dt^2): the slot algebra is not a
    # single stage in disguise
    eul, _ = Oracle(s).run(s.w, 0.0, dt, n, rk4=False)
    assert np.max(rel_err(data[: s.nve], eul)) > 1e-8


def test_driver_rk4calls_stratified_matches_fused(tmp_path):
    # gravity, the vertical driver at the stage times and contcd4 ghosts of
    # every slot; the fused branch on the same input gives the same state
    s = _strat_small()
    s.w = Oracle(s).boundary(perturbed(s, amp=1e-3, seed=12))
    inp = roundtrip_state(s, tmp_path, "strat_rkc")
    n, t0 = 3, 1.0
    outs = {}
    for flag in ("-rk4calls", "-rk4"):
        outs[flag] = str(tmp_path / f"strat_rkc_out{flag}.ini")
        run_driver(3, ["-i", inp, "-o", outs[flag], "-n", str(n), "-dt", repr(s.dt), "-t0",
                       str(t0), "-bc", "4,4,4", "-driver", _drv_arg(s), flag], tmp_path)
    _, _, calls = vacio.read_ascii(outs["-rk4calls"])
    _, _, fused = vacio.read_ascii(outs["-rk4"])
    ref, _ = Oracle(s).run(s.w, t0, s.dt, n, rk4=True)
    err = rel_err(calls[: s.nve], ref)
    print("driver rk4calls stratified rel err", err)
    assert np.all(err < 1e-11), err
    assert np.all(rel_err(calls[: s.nve], fused[: s.nve]) < 1e-11)


def test_driver_torsional_onestep_and_rk4(tmp_path):
    # the C driver with the reference's torsional user source (-torsional):
    # per-entry-point onestep (cusource with p->qt) and fused RK4
    s = configs.torsional_driver(_strat_small())
    o = Oracle(s)
    s.w = o.boundary(perturbed(s, amp=1e-3, seed=10))
    inp = roundtrip_state(s, tmp_path, "strat_tor")
    n, t0 = 3, 5.0
    for rk4 in (False, True):
        out = str(tmp_path / f"strat_tor_out{int(rk4)}.ini")
        args = ["-i", inp, "-o", out, "-n", str(n), "-dt", repr(s.dt), "-t0", str(t0),
                "-bc", "4,4,4", "-driver", _drv_arg(s), "-torsional"] + (["-rk4"] if rk4 else [])
        run_driver(3, args, tmp_path)
        _, _, data = vacio.read_ascii(out)
        if rk4:
            ref, _ = Oracle(s).run(s.w, t0, s.dt, n, rk4=True)
        else:  # the driver advances p->qt before each step (iosmaug.cpp:918)
            ref, _ = Oracle(s).run(s.w, t0 + s.dt, s.dt, n, rk4=False)
        err = rel_err(data[: s.nve], ref)
        print("driver torsional rk4" if rk4 else "driver torsional onestep", "rel err", err)
        assert np.all(err < 1e-11), err


def test_driver_onestep_reference_bw_ini(tmp_path):
    """the reference's own Brio-Wu initial state (configs/zero1_BW_asc.ini,
    800 x 6 cells, committed gzipped) stepped by the drop-in driver against
    the oracle.  Its y axis (two interior cells) is widened to the 8 cells per
    axis the path needs by repeating its rows: the state is the same in every
    row, and at 6 cells a contcd4 ghost's image (n-5 = 1) would itself be a
    ghost cell."""
    import gzip
    import shutil
    ini = str(tmp_path / "zero1_BW_asc.ini")
    with gzip.open(os.path.join(ROOT, "tests", "golden", "zero1_BW_asc.ini.gz"), "rb") as fi, \
            open(ini, "wb") as fo:
        shutil.copyfileobj(fi, fo)
    hdr, pos, data = vacio.read_ascii(ini)
    n0, n1 = hdr["dims"]
    row = data[:, :n0]
    assert np.array_equal(data, np.tile(row, (1, n1)))  # the same in every row
    ny = 8
    s = configs.brio_wu(ncells=n0 - 4, ny=ny - 4)
    s.gamma = hdr["eqpar"][0]
    s.w = np.tile(row[:6], (1, ny))
    s.bg = np.tile(row[6:], (1, ny))
    x = pos[0][:n0]
    s.pos[0] = x
    s.dx[0] = np.full(n0, x[1] - x[0])
    inp = roundtrip_state(s, tmp_path, "bw_ref")
    out = str(tmp_path / "bw_ref_out.ini")
    dt, n = 1e-4, 10
    run_driver(2, ["-i", inp, "-o", out, "-n", str(n), "-dt", repr(dt), "-bc", "4,4"], tmp_path)
    _, _, res = vacio.read_ascii(out)
    ref, _ = Oracle(s).run(s.w, dt, dt, n, rk4=False)
    err = rel_err(res[: s.nve], ref)
    print("driver onestep on the reference's Brio-Wu file: rel err", err)
    assert np.all(err < 1e-11), err
    # the waves have left the interface
    assert np.abs(ref[0] - s.w[0]).max() > 1e-3


@pytest.mark.parametrize("bc", [(3, 5), (5, 3), (3, 4)], ids=["cont_fixed", "fixed_cont",
                                                          "cont_contcd4"])
def test_driver_onestep_cont_and_fixed_2d(tmp_path, bc):
    """cont (3) and fixed (5) faces through the drop-in driver: the driver
    records the fixed faces with the p->it == -1 call before the loop
    (iosmaug.cpp:645-659); the oracle keeps the ghost values of the state as
    read from the file"""
    s = configs.brio_wu(ncells=60, ny=12)
    s.bc = bc
    s.w = perturbed(s, amp=1e-3, seed=40)
    inp = roundtrip_state(s, tmp_path, "bw_bc")
    out = str(tmp_path / "bw_bc_out.ini")
    dt, n = 5e-4, 6
    run_driver(2, ["-i", inp, "-o", out, "-n", str(n), "-dt", repr(dt),
                   "-bc", ",".join(str(b) for b in bc)], tmp_path)
    _, _, res = vacio.read_ascii(out)
    o = Oracle(s)
    o.set_fixed(s.w)
    ref, _ = o.run(s.w, dt, dt, n, rk4=False)
    err = rel_err(res[: s.nve], ref)
    print(f"driver onestep bc={bc} rel err", err)
    assert np.all(err < 1e-11), err
    if 5 in bc:
        # the fixed faces still hold the values read from the file (the
        # corners of dim 0's faces are refilled by dim 1, filled after it)
        d = bc.index(5)
        i = np.tile(np.arange(s.n[0]), s.n[1])
        j = np.repeat(np.arange(s.n[1]), s.n[0])
        x, y = (i, j) if d == 0 else (j, i)
        gh = (x < 2) | (x > s.n[d] - 3)
        if d == 0:
            gh &= (y >= 2) & (y <= s.n[1] - 3)
        assert np.array_equal(res[: s.nve][:, gh], s.w[:, gh])


def test_driver_onestep_fixed_vertical_stratified(tmp_path):
    # the stratified atmosphere with its bottom and top faces fixed
    s = _strat_small()
    s.bc = (5, 4, 4)
    o = Oracle(s)
    s.w = perturbed(s, amp=1e-3, seed=41)
    inp = roundtrip_state(s, tmp_path, "strat_fix")
    out = str(tmp_path / "strat_fix_out.ini")
    n, t0 = 3, 1.5
    run_driver(3, ["-i", inp, "-o", out, "-n", str(n), "-dt", repr(s.dt), "-t0", str(t0),
                   "-bc", "5,4,4", "-driver", _drv_arg(s)], tmp_path)
    _, _, res = vacio.read_ascii(out)
    o = Oracle(s)
    o.set_fixed(s.w)
    ref, _ = o.run(s.w, t0 + s.dt, s.dt, n, rk4=False)
    err = rel_err(res[: s.nve], ref)
    print("driver onestep fixed vertical faces rel err", err)
    assert np.all(err < 1e-11), err
