"""The C-ABI libraries load on a host without a GPU and export every function
that include/*.h declares.  No compute call is made here."""
import ctypes
import os
import re
import subprocess

import pytest

from conftest import ROOT as REPO

INC = os.path.join(REPO, "include")
LIB = os.path.join(REPO, "smaug-all_amd", "lib")

DECL = re.compile(r"^(?:const\s+)?(?:int|long long|char|void|double)\s*\**\s*(\w+)\s*\(", re.M)


def declared(header):
    with open(os.path.join(INC, header)) as f:
        text = f.read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(DECL.findall(text)))


def exported(lib):
    out = subprocess.run(["nm", "-D", "--defined-only", lib], capture_output=True, text=True,
                         check=True).stdout
    return {line.split()[-1] for line in out.splitlines() if " T " in line}


def test_engine_header_parsed():
    names = declared("smaug_engine.h")
    for n in ("smaug_create", "smaug_rk4", "smaug_ens_create", "smaug_ens_rk4"):
        assert n in names


@pytest.mark.parametrize("lib,header", [
    ("libsmaug_engine.so", "smaug_engine.h"),
    ("libsmaug_hip2d.so", "smaug_hip.h"),
    ("libsmaug_hip3d.so", "smaug_hip.h"),
])
def test_library_exports_header(lib, header):
    path = os.path.join(LIB, lib)
    assert os.path.exists(path), f"{path} not built"
    ctypes.CDLL(path)  # resolves its HIP runtime dependencies
    names = declared(header)
    assert len(names) > 10
    missing = [n for n in names if n not in exported(path)]
    assert not missing, f"{lib} lacks {missing}"


def test_python_binding_lists_every_engine_symbol():
    import smaugamd.engine as eng
    assert sorted(eng.EXPORTS) == declared("smaug_engine.h")


def test_grid_below_minimum_rejected_before_device_use():
    # smaug_create checks the configuration before it looks for a device, so
    # the size error is the one reported here as on a GPU host
    import sys
    sys.path.insert(0, os.path.join(REPO, "smaug-all_amd"))
    from smaugamd import configs
    from smaugamd.engine import FP64, Engine, SmaugError
    s = configs.stratified(nz=3, nx=4, ny=4)   # 7 cells along the vertical axis
    with pytest.raises(SmaugError, match=r"n\[0\]=7: need >= 8"):
        Engine(s, FP64)


def test_sums_noalloc_rejects_bad_sizes_before_device_use():
    # argument and size checks come before any HIP call
    from smaugamd import engine as eng
    big = (2048, 2048, 2048)  # 2^33 cells: above the 2^29-cell limit
    assert eng.sums_noalloc(3, big, eng.MIXED, 1, 1) == 1
    assert "exceed the limit" in eng.last_error()
    assert eng.sums_noalloc(3, (4, 8, 8), eng.FP64, 1, 1) == 1
    assert "n[0]=4" in eng.last_error()
    assert eng.sums_noalloc(4, (8, 8, 8), eng.FP64, 1, 1) == 1
    assert eng.sums_noalloc(3, (8, 8, 8), 7, 1, 1) == 1
    assert eng.sums_noalloc(3, (8, 8, 8), eng.FP64, 0, 1) == 1


@pytest.mark.parametrize("extra", [["-mixed"], ["-moddton"]])
def test_driver_rk4calls_rejects_mixed_and_moddton(extra):
    # the per-entry-point RK4 loop runs in FP64 with a fixed dt; the option
    # check comes before the configuration is read or a device is touched
    exe = os.path.join(LIB, "smaug_main2d")
    r = subprocess.run([exe, "-i", "/nonexistent.ini", "-rk4calls"] + extra, capture_output=True,
                       text=True, timeout=60)
    assert r.returncode == 2
    assert "-rk4calls runs in FP64 with a fixed dt" in r.stderr
